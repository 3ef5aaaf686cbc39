"""Seeded synthetic scenes written in the reference's on-disk format.

A scene directory holds ``scene.json`` (camera, hdri, pointLights as read by
SceneLoader.hpp:7-143), ``scene.mtl`` (Kd/Ks/map_Kd plus the additive Pr/Pm
keys), ``scene.obj`` (triangles, one ``o`` per mesh; the loader negates z,
ObjLoader.hpp:85-91, so world z is written negated) and optional ``.bmp``
albedo maps and ``HDRI/<name>.hdr``.

Workloads (sizes as stated in the project brief):
  sphere_field     64 UV spheres (32x16 segments, 960 tris) on a 2-tri ground
                   quad: 61,442 triangles, 65 materials, 2 point lights,
                   constant environment 0.5, pinhole 256x256.
  heightfield      1024x512-cell terrain (1,048,576 triangles) from 6-octave
                   value noise, 16 materials with 1024^2 albedo maps, 8 point
                   lights, 2048x1024 HDR sky with a sun lobe, 1920x1080 with
                   defocus (aperture 0.05 as the reference's f-number).
"""
from __future__ import annotations

import json
import math
import os
from pathlib import Path

import numpy as np


# --------------------------------------------------------------------------
# file writers
# --------------------------------------------------------------------------
def write_bmp(path, rgb8: np.ndarray):
    """24-bit BMP, bottom row first. rgb8: (H, W, 3) uint8, row 0 = top."""
    h, w, _ = rgb8.shape
    stride = (w * 3 + 3) & ~3
    data = np.zeros((h, stride), np.uint8)
    bgr = rgb8[::-1, :, ::-1].reshape(h, w * 3)
    data[:, :w * 3] = bgr
    size = 54 + data.size
    hdr = bytearray(54)
    hdr[0:2] = b"BM"
    hdr[2:6] = size.to_bytes(4, "little")
    hdr[10:14] = (54).to_bytes(4, "little")
    hdr[14:18] = (40).to_bytes(4, "little")
    hdr[18:22] = w.to_bytes(4, "little")
    hdr[22:26] = h.to_bytes(4, "little")
    hdr[26:28] = (1).to_bytes(2, "little")
    hdr[28:30] = (24).to_bytes(2, "little")
    hdr[34:38] = data.size.to_bytes(4, "little")
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(data.tobytes())


def rgbe_encode(rgb: np.ndarray) -> np.ndarray:
    """Float RGB -> RGBE bytes (Ward's shared-exponent format)."""
    m = rgb.max(axis=-1)
    out = np.zeros(rgb.shape[:-1] + (4,), np.uint8)
    nz = m > 1e-32
    mant, ex = np.frexp(m[nz])
    scale = mant * 256.0 / m[nz]
    out[nz, 0:3] = np.clip(rgb[nz] * scale[:, None], 0, 255).astype(np.uint8)
    out[nz, 3] = (ex + 128).astype(np.uint8)
    return out


def write_hdr(path, rgb: np.ndarray):
    """Radiance .hdr, -Y H +X W, new-style RLE scanlines (literal runs)."""
    h, w, _ = rgb.shape
    e = rgbe_encode(rgb.astype(np.float32))
    head = f"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y {h} +X {w}\n".encode()
    chunks = [head]
    nfull, rem = divmod(w, 128)
    for y in range(h):
        line = [bytes([2, 2, (w >> 8) & 0xFF, w & 0xFF])]
        for c in range(4):
            comp = e[y, :, c]
            for k in range(nfull):
                line.append(bytes([128]) + comp[k * 128:(k + 1) * 128].tobytes())
            if rem:
                line.append(bytes([rem]) + comp[nfull * 128:].tobytes())
        chunks.append(b"".join(line))
    with open(path, "wb") as f:
        f.write(b"".join(chunks))


def _fmt_rows(prefix: str, a: np.ndarray, fmt: str) -> str:
    if a.size == 0:
        return ""
    row = prefix + " " + " ".join([fmt] * a.shape[1]) + "\n"
    return (row * a.shape[0]) % tuple(a.ravel().tolist())


class ObjWriter:
    """Accumulates meshes in world coordinates and writes obj/mtl text."""

    def __init__(self):
        self.parts = []
        self.nv = self.nvt = self.nvn = 0

    def add_mesh(self, name, material, verts, uvs, normals, faces):
        """faces: (F, 3) indices into this mesh's verts/uvs/normals (same index)."""
        v = np.asarray(verts, np.float64).copy()
        n = np.asarray(normals, np.float64).copy()
        v[:, 2] *= -1   # the loader negates z
        n[:, 2] *= -1
        f = np.asarray(faces, np.int64) + 1
        fv, fvt, fvn = f + self.nv, f + self.nvt, f + self.nvn
        s = [f"o {name}\n", _fmt_rows("v", v, "%.6f"), _fmt_rows("vt", np.asarray(uvs, np.float64), "%.6f"),
             _fmt_rows("vn", n, "%.6f"), f"usemtl {material}\n"]
        tri = np.stack([fv, fvt, fvn], axis=-1).reshape(-1, 9)
        s.append(("f %d/%d/%d %d/%d/%d %d/%d/%d\n" * tri.shape[0]) % tuple(tri.ravel().tolist()))
        self.parts.append("".join(s))
        self.nv += len(v)
        self.nvt += len(uvs)
        self.nvn += len(n)

    def write(self, path):
        with open(path, "w") as fh:
            fh.write("# synthetic scene\n")
            for p in self.parts:
                fh.write(p)


def write_mtl(path, mats):
    """mats: list of dicts name, Kd, Ks, Pr, Pm, map_Kd (optional)."""
    out = []
    for m in mats:
        out.append(f"newmtl {m['name']}\n")
        kd = m.get("Kd", (0.5, 0.5, 0.5))
        out.append("Kd %.6f %.6f %.6f\n" % tuple(kd))
        ks = m.get("Ks", (0.5, 0.5, 0.5))
        out.append("Ks %.6f %.6f %.6f\n" % tuple(ks))
        if "Pr" in m:
            out.append("Pr %.6f\n" % m["Pr"])
        if "Pm" in m:
            out.append("Pm %.6f\n" % m["Pm"])
        if "map_Kd" in m:
            out.append(f"map_Kd {m['map_Kd']}\n")
        out.append("\n")
    Path(path).write_text("".join(out))


def write_json(path, camera, hdri, lights):
    doc = {"camera": camera, "hdri": hdri,
           "pointLights": [{"position": {"x": p[0], "y": p[1], "z": p[2]},
                            "radiance": {"x": r[0], "y": r[1], "z": r[2]}} for p, r in lights]}
    Path(path).write_text(json.dumps(doc, indent=1))


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
def uv_sphere(center, radius, seg=32, rings=16):
    """UV sphere with pole fans: seg*(rings-2)*2 + 2*seg triangles (960 for 32x16)."""
    verts, uvs, nrms = [], [], []
    # poles are split per segment so every fan triangle has its own uv
    for r in range(rings + 1):
        th = math.pi * r / rings
        for s in range(seg + 1):
            ph = 2 * math.pi * s / seg
            n = (math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph))
            nrms.append(n)
            verts.append((center[0] + radius * n[0], center[1] + radius * n[1], center[2] + radius * n[2]))
            uvs.append((s / seg, 1 - r / rings))
    idx = lambda r, s: r * (seg + 1) + s
    faces = []
    for r in range(rings):
        for s in range(seg):
            a, b, c, d = idx(r, s), idx(r, s + 1), idx(r + 1, s), idx(r + 1, s + 1)
            if r == 0:
                faces.append((a, d, c))
            elif r == rings - 1:
                faces.append((a, b, c))
            else:
                faces.append((a, b, d))
                faces.append((a, d, c))
    return np.array(verts), np.array(uvs), np.array(nrms), np.array(faces)


def checker_noise_texture(size, seed):
    """size x size 8-bit RGB: an 8x8 checker of two seeded colours modulated by value noise."""
    rng = np.random.default_rng(seed)
    c0, c1 = rng.uniform(0.2, 0.9, 3), rng.uniform(0.2, 0.9, 3)
    y, x = np.mgrid[0:size, 0:size] / size
    chk = ((np.floor(x * 8) + np.floor(y * 8)) % 2)[..., None]
    n = value_noise(x * 8, y * 8, seed, octaves=4)[..., None]
    rgb = (c0 * (1 - chk) + c1 * chk) * (0.75 + 0.25 * n)
    return (np.clip(rgb, 0, 1) * 255).astype(np.uint8)


def sphere_field(out_dir, n_spheres=64, res=(256, 256), seed=0, seg=32, rings=16, ground_texture=None,
                 env=0.5, lights=2, light_radiance=50.0, extent=10.0, hdri_sky=None, ground_tex_size=0):
    """Sphere-field scene (the 61,442-triangle workload at n_spheres=64).
    ground_tex_size > 0: the ground gets a seeded checker/noise albedo map of
    that size (ground_texture gives one explicitly)."""
    if ground_texture is None and ground_tex_size > 0:
        ground_texture = checker_noise_texture(ground_tex_size, seed + 1000)
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    ow = ObjWriter()
    mats = []
    g = 20.0
    gv = np.array([(-g, 0, -g), (g, 0, -g), (g, 0, g), (-g, 0, g)], np.float64)
    guv = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float64)
    gn = np.tile([0.0, 1.0, 0.0], (4, 1))
    ow.add_mesh("ground", "mat_ground", gv, guv, gn, np.array([(0, 1, 2), (0, 2, 3)]))
    gm = {"name": "mat_ground", "Kd": (0.5, 0.5, 0.5), "Ks": (0.5, 0.5, 0.5), "Pr": 1.0, "Pm": 0.0}
    if ground_texture is not None:
        write_bmp(out / "ground.bmp", ground_texture)
        gm["map_Kd"] = "ground.bmp"   # resolved against the scene directory
    mats.append(gm)
    centers = rng.uniform(-extent, extent, size=(n_spheres, 2))
    kd = rng.uniform(0.1, 0.9, size=(n_spheres, 3))
    pr = rng.uniform(0.05, 1.0, size=n_spheres)
    pm = (rng.uniform(size=n_spheres) < 0.3).astype(np.float64)
    for i in range(n_spheres):
        v, uv, nr, f = uv_sphere((centers[i, 0], 1.0, centers[i, 1]), 1.0, seg, rings)
        ow.add_mesh(f"sphere_{i}", f"mat_{i}", v, uv, nr, f)
        mats.append({"name": f"mat_{i}", "Kd": tuple(kd[i]), "Ks": (0.5, 0.5, 0.5), "Pr": float(pr[i]),
                     "Pm": float(pm[i])})
    ow.write(out / "scene.obj")
    write_mtl(out / "scene.mtl", mats)
    lp = [(float(rng.uniform(-extent, extent)), float(rng.uniform(5, 10)), float(rng.uniform(-extent, extent)))
          for _ in range(lights)]
    pitch = math.degrees(math.atan2(8.0, 25.0))
    camera = {"xRes": res[0], "yRes": res[1], "focalLength": 0.035, "focusDistance": 26.0, "aperture": 2.8,
              "bokeh": False, "position": {"x": 0.0, "y": 8.0, "z": -25.0},
              "rotation": {"x": pitch, "y": 0.0, "z": 0.0}}
    if hdri_sky is not None:
        (out / "HDRI").mkdir(exist_ok=True)
        write_hdr(out / "HDRI" / "sky.hdr", hdri_sky)
        hdri = {"name": "sky"}
    else:
        hdri = {"color": {"r": env, "g": env, "b": env}}
    write_json(out / "scene.json", camera, hdri, [(p, (light_radiance,) * 3) for p in lp])
    return out


def sphere_field_arrays(n_spheres=64, res=(256, 256), seed=0, seg=32, rings=16, hdri=None, env=0.5, lights=2,
                        light_radiance=50.0, extent=10.0):
    """The sphere field of sphere_field(), built in memory as a SceneData (no
    files): world-space triangles, the same material parameters (albedo ~
    U[0.1,0.9]^3, roughness ~ U[0.05,1], metallic ~ Bernoulli(0.3), specular
    0.5), lights and camera.  `hdri`: float RGB array or None for a constant
    environment `env`."""
    from . import SceneData, make_material
    rng = np.random.default_rng(seed)
    g = 20.0
    verts, nrms, uvs, oid = [], [], [], []
    gv = np.array([(-g, 0, -g), (g, 0, -g), (g, 0, g), (-g, 0, g)], np.float64)
    guv = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float64)
    for f in ((0, 1, 2), (0, 2, 3)):
        verts.append(gv[list(f)].reshape(9))
        nrms.append(np.tile([0.0, 1.0, 0.0], 3))
        uvs.append(guv[list(f)].reshape(6))
        oid.append(0)
    mats = [make_material(albedo=(0.5, 0.5, 0.5), roughness=1.0, metallic=0.0, specular=0.5)]
    centers = rng.uniform(-extent, extent, size=(n_spheres, 2))
    kd = rng.uniform(0.1, 0.9, size=(n_spheres, 3))
    pr = rng.uniform(0.05, 1.0, size=n_spheres)
    pm = (rng.uniform(size=n_spheres) < 0.3).astype(np.float64)
    V, UV, N, F = uv_sphere((0.0, 0.0, 0.0), 1.0, seg, rings)
    tv, tn, tu = V[F].reshape(-1, 9), N[F].reshape(-1, 9), UV[F].reshape(-1, 6)
    for i in range(n_spheres):
        off = np.tile([centers[i, 0], 1.0, centers[i, 1]], 3)
        verts.append(tv + off)
        nrms.append(tn)
        uvs.append(tu)
        oid.append(np.full(len(tv), i + 1))
        mats.append(make_material(albedo=tuple(kd[i]), roughness=float(pr[i]), metallic=float(pm[i]), specular=0.5))
    v = np.vstack([np.atleast_2d(x) for x in verts]).astype(np.float32)
    n = np.vstack([np.atleast_2d(x) for x in nrms]).astype(np.float32)
    uv = np.vstack([np.atleast_2d(x) for x in uvs]).astype(np.float32)
    ob = np.concatenate([np.atleast_1d(x) for x in oid]).astype(np.int32)
    lp = [((float(rng.uniform(-extent, extent)), float(rng.uniform(5, 10)), float(rng.uniform(-extent, extent))),
           (light_radiance,) * 3) for _ in range(lights)]
    pitch = math.degrees(math.atan2(8.0, 25.0))
    cam = dict(x_res=res[0], y_res=res[1], position=(0.0, 8.0, -25.0), rotation=(pitch, 0.0, 0.0),
               focus_distance=26.0)
    env_img = hdri if hdri is not None else np.full((4, 8, 3), env, np.float32)
    return SceneData.from_arrays(v, cam, normals=n, uvs=uv, object_id=ob, mesh_material=np.arange(len(mats)),
                                 materials=mats, hdri=env_img, lights=lp)


def batch_workload(n_scenes=32, seed=0, tri_range=(1000, 500000), side_range=(128, 1024), spp_range=(4, 64)):
    """Seeded batch of sphere fields: triangle counts log-uniform in tri_range
    (960 per sphere), square resolutions with side uniform in side_range,
    spp uniform in spp_range, each with its own 512x256 HDR sky (sun
    direction per scene).  Returns [(SceneData, spp, seed), ...]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_scenes):
        tris = math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))
        n_sph = max(1, int(round((tris - 2) / 960)))
        side = int(rng.integers(side_range[0], side_range[1] + 1))
        spp = int(rng.integers(spp_range[0], spp_range[1] + 1))
        sun = (float(rng.uniform(-1, 1)), float(rng.uniform(0.3, 1.0)), float(rng.uniform(-1, 1)))
        sky = sky_hdr(512, 256, sun_dir=sun, peak=200.0, sigma_deg=3.0)
        ext = 10.0 * max(1.0, math.sqrt(n_sph / 64.0))
        sc = sphere_field_arrays(n_spheres=n_sph, res=(side, side), seed=seed * 1000 + i, hdri=sky, extent=ext)
        out.append((sc, spp, seed * 1000 + i))
    return out


def value_noise(x, z, seed, octaves=6):
    """sum_o 2^-o * valuenoise(2^o p) with a seeded lattice and smoothstep."""
    rng = np.random.default_rng(seed)
    lat = rng.uniform(-1, 1, size=(256, 256))
    tot = np.zeros_like(x)
    for o in range(octaves):
        f = 2.0 ** o
        px, pz = x * f, z * f
        ix, iz = np.floor(px).astype(np.int64), np.floor(pz).astype(np.int64)
        fx, fz = px - ix, pz - iz
        sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
        g = lambda a, b: lat[a & 255, b & 255]
        v = (g(ix, iz) * (1 - sx) * (1 - sz) + g(ix + 1, iz) * sx * (1 - sz) + g(ix, iz + 1) * (1 - sx) * sz +
             g(ix + 1, iz + 1) * sx * sz)
        tot += v * 2.0 ** (-o)
    return tot


def sky_hdr(w, h, sun_dir=(0.3, 0.6, 0.4), peak=1e3, sigma_deg=2.0):
    """Vertical gradient sky plus a Gaussian sun lobe (row 0 = zenith)."""
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    theta = v[:, None] * math.pi
    phi = u[None, :] * 2 * math.pi
    # direction of pixel (u, v) as the reference's reverseSphericalMapping, negated
    px, py, pz = np.cos(phi - math.pi), -np.cos(theta), -np.sin(phi - math.pi)
    a = np.sqrt(np.maximum(0, 1 - py * py))
    d = -np.stack([a * px + 0 * py, py + 0 * px, a * pz], axis=-1)
    sd = np.asarray(sun_dir, np.float64)
    sd = sd / np.linalg.norm(sd)
    ang = np.degrees(np.arccos(np.clip((d * sd).sum(-1), -1, 1)))
    up = np.clip(d[..., 1], -1, 1)
    base = np.stack([0.3 + 0.5 * np.clip(up, 0, 1), 0.4 + 0.5 * np.clip(up, 0, 1), 0.6 + 0.6 * np.clip(up, 0, 1)], -1)
    base[up < 0] *= 0.3
    sun = peak * np.exp(-0.5 * (ang / sigma_deg) ** 2)
    return (base + sun[..., None] * np.array([1.0, 0.95, 0.85])).astype(np.float32)


def heightfield(out_dir, cells=(1024, 512), extent=(51.2, 25.6), res=(1920, 1080), seed=0, n_mats=16,
                tex_size=1024, lights=8, sky=(2048, 1024), aperture=0.05, bokeh=True, textured=None):
    """Heightfield scene (the 1,048,576-triangle workload at the default size).
    `textured`: how many materials (the first ones) carry their albedo map;
    the others get a scalar albedo from a separate seeded stream (None: all
    textured; the geometry, maps, lights and camera do not depend on it)."""
    n_tex = n_mats if textured is None else textured
    rng_kd = np.random.default_rng(seed + 7)
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    cx, cz = cells
    xs = np.linspace(-extent[0], extent[0], cx + 1)
    zs = np.linspace(-extent[1], extent[1], cz + 1)
    X, Z = np.meshgrid(xs, zs)                        # (cz+1, cx+1)
    Y = 2.0 * value_noise(X / 8.0, Z / 8.0, seed + 1)
    # smooth vertex normals from central differences
    dx = np.gradient(Y, xs, axis=1)
    dz = np.gradient(Y, zs, axis=0)
    N = np.stack([-dx, np.ones_like(Y), -dz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    tiles = int(round(math.sqrt(n_mats)))
    tx, tz = cx // tiles, cz // tiles
    ow = ObjWriter()
    mats = []
    for k in range(n_mats):
        i, j = k % tiles, k // tiles
        c0, c1 = i * tx, (i + 1) * tx
        r0, r1 = j * tz, (j + 1) * tz
        vx = X[r0:r1 + 1, c0:c1 + 1]
        vz = Z[r0:r1 + 1, c0:c1 + 1]
        vy = Y[r0:r1 + 1, c0:c1 + 1]
        nn = N[r0:r1 + 1, c0:c1 + 1]
        verts = np.stack([vx, vy, vz], -1).reshape(-1, 3)
        uu = (np.arange(c1 - c0 + 1) / (c1 - c0))[None, :].repeat(r1 - r0 + 1, 0)
        vv = (np.arange(r1 - r0 + 1) / (r1 - r0))[:, None].repeat(c1 - c0 + 1, 1)
        uvs = np.stack([uu, vv], -1).reshape(-1, 2)
        W = c1 - c0 + 1
        a = (np.arange(r1 - r0)[:, None] * W + np.arange(c1 - c0)[None, :]).ravel()
        faces = np.concatenate([np.stack([a, a + 1, a + W + 1], -1), np.stack([a, a + W + 1, a + W], -1)])
        ow.add_mesh(f"tile_{k}", f"mat_{k}", verts, uvs, nn.reshape(-1, 3), faces)
        # albedo map: checker or value noise
        if k % 2 == 0:
            q = (np.arange(tex_size) // (tex_size // 16))
            chk = ((q[:, None] + q[None, :]) % 2).astype(np.float64)
            c_a, c_b = rng.uniform(0.1, 0.9, 3), rng.uniform(0.1, 0.9, 3)
            img = chk[..., None] * c_a + (1 - chk[..., None]) * c_b
        else:
            g = np.linspace(0, 16, tex_size)
            gx, gz = np.meshgrid(g, g)
            nz = 0.5 + 0.5 * value_noise(gx, gz, seed + 100 + k, octaves=4) / 2
            img = np.clip(nz, 0, 1)[..., None] * rng.uniform(0.3, 1.0, 3)
        name = f"albedo_{k}.bmp"
        write_bmp(out / name, (np.clip(img, 0, 1) * 255).astype(np.uint8))
        mat = {"name": f"mat_{k}", "Kd": (0.8, 0.8, 0.8), "Ks": (0.5, 0.5, 0.5),
               "Pr": float(rng.uniform(0.2, 1.0)), "Pm": float(rng.uniform() < 0.2), "map_Kd": name}
        if k >= n_tex:
            mat["Kd"] = tuple(float(x) for x in rng_kd.uniform(0.1, 0.9, 3))
            del mat["map_Kd"]
        mats.append(mat)
    ow.write(out / "scene.obj")
    write_mtl(out / "scene.mtl", mats)
    (out / "HDRI").mkdir(exist_ok=True)
    write_hdr(out / "HDRI" / "sky.hdr", sky_hdr(*sky))
    lp = [(float(rng.uniform(-extent[0], extent[0])), float(rng.uniform(6, 12)),
           float(rng.uniform(-extent[1], extent[1]))) for _ in range(lights)]
    # framed so the terrain fills the 16:9 view: the top row of pixels meets
    # the far edge, the bottom row the near half of the grid
    k = extent[0] / 51.2
    cam_pos = (0.0, 15.0 * k, -38.0 * k)
    pitch = 29.0
    focus = math.sqrt(cam_pos[1] ** 2 + cam_pos[2] ** 2)
    camera = {"xRes": res[0], "yRes": res[1], "focalLength": 0.035, "focusDistance": focus, "aperture": aperture,
              "bokeh": bokeh, "position": {"x": cam_pos[0], "y": cam_pos[1], "z": cam_pos[2]},
              "rotation": {"x": pitch, "y": 0.0, "z": 0.0}}
    write_json(out / "scene.json", camera, {"name": "sky"}, [(p, (60.0, 60.0, 60.0)) for p in lp])
    return out


def cached(kind: str, root=None, **kw) -> Path:
    """Generate a scene once per parameter set under `root` (default /tmp)."""
    root = Path(root or os.environ.get("ELEVEN_SCENE_CACHE", "/tmp/eleven_scenes"))
    key = kind + "_" + "_".join(f"{k}-{v}" for k, v in sorted(kw.items()) if not isinstance(v, np.ndarray))
    key = key.replace(" ", "").replace("(", "").replace(")", "").replace(",", "x")
    d = root / key
    if (d / "scene.json").exists():
        return d
    tmp = root / (key + f".tmp{os.getpid()}")
    {"sphere_field": sphere_field, "heightfield": heightfield}[kind](tmp, **kw)
    try:
        os.rename(tmp, d)
    except OSError:
        pass
    return d

"""The small scenes of the image-parity tests, regenerated from fixed seeds.

spheres          64 UV spheres at 64x64, constant environment 0.5
spheres_sky_tex  24 spheres at 64x48, textured ground, HDR sky with a sun lobe
terrain_bokeh    64x32-cell heightfield at 96x54, 4 textured materials,
                 4 point lights, HDR sky, defocus
"""
import numpy as np

import eleven
from eleven import scenegen

SCENES = ["spheres", "spheres_sky_tex", "terrain_bokeh"]
DEPTH = 5


def make_scene(kind, root):
    if kind == "spheres":
        d = scenegen.sphere_field(root / kind, res=(64, 64), seed=0)
    elif kind == "spheres_sky_tex":
        rng = np.random.default_rng(4)
        tex = (rng.uniform(0, 1, size=(32, 32, 3)) * 255).astype(np.uint8)
        d = scenegen.sphere_field(root / kind, n_spheres=24, res=(64, 48), seed=1, ground_texture=tex,
                                  hdri_sky=scenegen.sky_hdr(128, 64, peak=50.0, sigma_deg=6.0))
    elif kind == "terrain_bokeh":
        d = scenegen.heightfield(root / kind, cells=(64, 32), extent=(12.8, 6.4), res=(96, 54), seed=2,
                                 n_mats=4, tex_size=64, lights=4, sky=(128, 64), aperture=0.05, bokeh=True)
    else:
        raise ValueError(kind)
    return eleven.SceneData.load(d)


def maps_scene(res=(64, 48), seed=7, mirror_roughness=0.1):
    """Nine spheres on a ground quad, built in memory, exercising what the
    sphere fields leave out: a bilinear albedo map, a normal map (with corner
    tangents), roughness, metallic and emission maps, and the Disney lobes
    clearcoat, sheen (+ tint), subsurface, anisotropic and specular tint.
    HDR sky with a sun lobe and two point lights.  The metal sphere's
    roughness stays at 0.1 (0.0063 after the 2.2 power): at 0.05 (0.0014,
    next to GGX's 0.001 clamp) the roughness derivative of a mirror path
    cancels to a few FP32 ulps of a ~1e8 prefactor, and single-pixel adjoints
    then differ by ~10 % between any two FP32 evaluation orders (GPU, FP32
    oracle) and from the FP64 oracle."""
    import math
    rng = np.random.default_rng(seed)
    g = 8.0
    V, N, UV, T, OID = [], [], [], [], []
    gv = np.array([(-g, 0, -g), (g, 0, -g), (g, 0, g), (-g, 0, g)], np.float64)
    guv = np.array([(0, 0), (1, 0), (1, 1), (0, 1)], np.float64)
    for f in ((0, 1, 2), (0, 2, 3)):
        V.append(gv[list(f)].reshape(9))
        N.append(np.tile([0.0, 1.0, 0.0], 3))
        UV.append(guv[list(f)].reshape(6))
        T.append(np.tile([1.0, 0.0, 0.0], 3))
        OID.append(0)
    sv, suv, sn, sf = scenegen.uv_sphere((0.0, 0.0, 0.0), 1.0, 24, 12)
    # tangent along increasing longitude, (-z, 0, x) of the unit normal; (1, 0, 0) at the poles
    st = np.stack([-sn[:, 2], np.zeros(len(sn)), sn[:, 0]], -1)
    ln = np.linalg.norm(st, axis=-1, keepdims=True)
    st = np.where(ln > 1e-6, st / np.maximum(ln, 1e-12), np.array([1.0, 0.0, 0.0]))
    for i in range(9):
        c = np.array([-4.0 + 4.0 * (i % 3), 1.0, -3.0 + 3.0 * (i // 3)])
        V.append((sv[sf] + c).reshape(-1, 9))
        N.append(sn[sf].reshape(-1, 9))
        UV.append(suv[sf].reshape(-1, 6))
        T.append(st[sf].reshape(-1, 9))
        OID.append(np.full(len(sf), i + 1))
    v = np.vstack([np.atleast_2d(x) for x in V]).astype(np.float32)
    n = np.vstack([np.atleast_2d(x) for x in N]).astype(np.float32)
    uv = np.vstack([np.atleast_2d(x) for x in UV]).astype(np.float32)
    tg = np.vstack([np.atleast_2d(x) for x in T]).astype(np.float32)
    ob = np.concatenate([np.atleast_1d(x) for x in OID]).astype(np.int32)
    albedo = rng.uniform(0.05, 0.95, size=(32, 32, 3)).astype(np.float32)
    nmap = np.concatenate([0.5 + 0.25 * rng.uniform(-1, 1, size=(16, 16, 2)), np.ones((16, 16, 1))], -1)
    rough = np.repeat(rng.uniform(0.1, 1.0, size=(16, 16, 1)), 3, -1)
    emis = rng.uniform(0.0, 2.0, size=(8, 8, 3))
    metal = np.repeat(rng.uniform(0.0, 1.0, size=(16, 16, 1)), 3, -1)
    textures = [eleven.make_texture(albedo, filt=1), eleven.make_texture(nmap.astype(np.float32)),
                eleven.make_texture(rough.astype(np.float32)), eleven.make_texture(emis.astype(np.float32)),
                eleven.make_texture(metal.astype(np.float32), filt=1)]
    mk = eleven.make_material
    mats = [mk(albedo=(0.6, 0.6, 0.6), roughness=0.5, albedo_tex=0, normal_tex=1, roughness_tex=2),
            mk(albedo=(0.8, 0.3, 0.2), roughness=0.4, clearcoat=0.8, clearcoat_gloss=0.6),
            mk(albedo=(0.2, 0.7, 0.3), roughness=0.9, sheen=0.7, sheen_tint=0.5),
            mk(albedo=(0.3, 0.4, 0.9), roughness=0.6, subsurface=0.8),
            mk(albedo=(0.9, 0.8, 0.3), roughness=0.3, metallic=0.7, anisotropic=0.8),
            mk(albedo=(0.7, 0.2, 0.7), roughness=0.5, specular=0.9, specular_tint=0.7),
            mk(albedo=(0.5, 0.5, 0.5), roughness=0.7, emission_tex=3),
            mk(albedo=(0.6, 0.5, 0.4), roughness=0.35, metallic_tex=4),
            mk(albedo=(0.4, 0.6, 0.6), roughness=0.25, normal_tex=1, clearcoat=0.4, sheen=0.3, subsurface=0.3),
            mk(albedo=(0.9, 0.9, 0.9), roughness=mirror_roughness, metallic=1.0)]
    lights = [((3.0, 6.0, -4.0), (30.0, 30.0, 30.0)), ((-5.0, 4.0, 2.0), (15.0, 20.0, 25.0))]
    pitch = math.degrees(math.atan2(5.0, 12.0))
    cam = dict(x_res=res[0], y_res=res[1], position=(0.0, 5.0, -12.0), rotation=(pitch, 0.0, 0.0))
    return eleven.SceneData.from_arrays(v, cam, normals=n, uvs=uv, tangents=tg, object_id=ob,
                                        mesh_material=np.arange(len(mats)), materials=mats, textures=textures,
                                        hdri=scenegen.sky_hdr(64, 32, peak=40.0, sigma_deg=8.0), lights=lights)


def deep_stack_scene(n=20000, res=(64, 64), seed=31):
    """A cloud of large, heavily overlapping triangles around the camera: most
    rays enter many sibling boxes at every level, so traversal stacks grow
    past the LDS part and spill to the global area (the tree's stack bound
    is far above 16).  Face normals, one diffuse material, a sky gradient and
    one point light, so both shadow launches of every bounce have rays."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-3.0, 3.0, size=(n, 1, 3))
    v = c + rng.normal(scale=0.8, size=(n, 3, 3))
    fn = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    fn /= np.maximum(np.linalg.norm(fn, axis=-1, keepdims=True), 1e-12)
    normals = np.repeat(fn[:, None, :], 3, axis=1)
    sky = np.zeros((16, 32, 3), np.float32)
    sky[...] = np.linspace(1.5, 0.2, 16, dtype=np.float32)[:, None, None]
    return eleven.SceneData.from_arrays(v.astype(np.float32).reshape(n, 9), camera=dict(x_res=res[0], y_res=res[1]),
                                        normals=normals.astype(np.float32).reshape(n, 9), hdri=sky,
                                        lights=[((0.5, 4.0, 0.5), (20.0, 20.0, 20.0))])

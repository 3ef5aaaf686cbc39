"""Regression tests for defects found by review, each run on the fixed code.

1. Paths deeper than 32 bounces with zero-throughput culling: the wavefront
   renderer's per-path mask of NaN-prone bounces covers every bounce up to
   the 64 el_render accepts, so culled and faithful renders, plain and
   wavefront, keep identical accepted-sample counts at depth 40.  The scene
   is a closed mirror-like box (GGX reflections below the horizon zero the
   throughput: kernel.cu:404, Disney.hpp:124-125, 200) with a point light
   inside and a sky whose top row (pdf infinite, HDRI.hpp:151) holds about
   1 % of the luminance, so samples that draw it at any bounce are rejected
   as NaN (kernel.cu:449).
2. el_set_camera discards the device residual of el_residual_l2: an adjoint
   that asks for it afterwards fails instead of using the previous view's.
3. A batch rendered with a captured graph and then with EL_BATCH_NO_CULL
   replays the faithful launch sequence, not the culling one.
4. The wavefront renderer's per-path AOV buffer grows with the generation:
   a context whose other wavefront buffers a large non-AOV render has grown,
   and whose AOV buffer a small AOV render has sized, renders a larger AOV
   frame exactly as a fresh context does (the buffer was once sized by the
   first AOV render only, and a larger one then wrote past its end).
"""
import numpy as np
import pytest

import eleven
from eleven import scenegen


def _box_scene(res=(24, 24), sky_h=64):
    """A closed box [-2,2]^3 seen from its centre, walls facing inward."""
    c = 2.0
    P = np.array([[x, y, z] for x in (-c, c) for y in (-c, c) for z in (-c, c)], np.float64)
    idx = lambda x, y, z: 4 * x + 2 * y + z
    quads = [  # (corner indices in order, inward normal)
        ([idx(0, 0, 0), idx(0, 1, 0), idx(0, 1, 1), idx(0, 0, 1)], (1, 0, 0)),
        ([idx(1, 0, 0), idx(1, 0, 1), idx(1, 1, 1), idx(1, 1, 0)], (-1, 0, 0)),
        ([idx(0, 0, 0), idx(0, 0, 1), idx(1, 0, 1), idx(1, 0, 0)], (0, 1, 0)),
        ([idx(0, 1, 0), idx(1, 1, 0), idx(1, 1, 1), idx(0, 1, 1)], (0, -1, 0)),
        ([idx(0, 0, 0), idx(1, 0, 0), idx(1, 1, 0), idx(0, 1, 0)], (0, 0, 1)),
        ([idx(0, 0, 1), idx(0, 1, 1), idx(1, 1, 1), idx(1, 0, 1)], (0, 0, -1)),
    ]
    V, N, OID = [], [], []
    for k, (q, n) in enumerate(quads):
        for tri in ((q[0], q[1], q[2]), (q[0], q[2], q[3])):
            V.append(P[list(tri)].reshape(9))
            N.append(np.tile(n, 3))
            OID.append(k % 2)
    mats = [eleven.make_material(albedo=(0.9, 0.8, 0.7), roughness=0.6, metallic=1.0),
            eleven.make_material(albedo=(0.7, 0.8, 0.9), roughness=0.4, metallic=0.6)]
    sky = np.full((sky_h, 2 * sky_h, 3), 0.5, np.float32)
    # the pole row (pdf = lum / sum * W * H / (2 pi sin 0) = inf) holds ~1 %
    # of the luminance: about one bounce in a hundred draws it
    sky[0] = 0.3
    cam = dict(x_res=res[0], y_res=res[1], position=(0.0, 0.0, -1.5), rotation=(10.0, 5.0, 0.0))
    return eleven.SceneData.from_arrays(np.array(V, np.float32), cam, normals=np.array(N, np.float32),
                                        object_id=np.array(OID, np.int32), mesh_material=np.array([0, 1]),
                                        materials=mats, hdri=sky, lights=[((0.5, 1.0, 0.5), (20.0, 20.0, 20.0))])


@pytest.mark.gpu
def test_depth_40_culling_keeps_counts(gpu_ctx):
    print("[regress] depth 40: culled vs faithful, plain vs wavefront", flush=True)
    s = _box_scene()
    gpu_ctx.upload(s)
    out = {}
    for name, mode, flags in [("wf_cull", eleven.MODE_WAVEFRONT, 0),
                              ("wf_faithful", eleven.MODE_WAVEFRONT, eleven.FLAG_NO_CULL),
                              ("plain_cull", eleven.MODE_PLAIN, 0)]:
        gpu_ctx.clear()
        st = gpu_ctx.render(8, depth=40, seed=17, mode=mode, flags=flags)
        img, cnt, _, _ = gpu_ctx.read()
        out[name] = (img[..., :3].astype(np.float64), cnt, st.rays)
    ref_img, ref_cnt, faithful_rays = out["wf_faithful"]
    # the test is meaningful: samples are rejected, and culling stops paths
    assert ref_cnt.sum() < 8 * ref_cnt.size and ref_cnt.sum() > 0
    assert out["wf_cull"][2] < faithful_rays
    for name in ("wf_cull", "plain_cull"):
        img, cnt, _ = out[name]
        assert np.array_equal(cnt, ref_cnt), f"{name}: accepted-sample counts differ from the faithful render"
        np.testing.assert_allclose(img, ref_img, rtol=1e-4, atol=1e-6)


@pytest.mark.gpu
def test_set_camera_discards_device_residual(gpu_ctx):
    print("[regress] set_camera drops the device residual", flush=True)
    s = scenegen.sphere_field_arrays(n_spheres=6, res=(16, 16), seed=4)
    gpu_ctx.upload(s)
    gpu_ctx.render(2, depth=3, seed=1)
    gpu_ctx.residual_l2(np.zeros((16, 16, 3), np.float32))
    gpu_ctx.adjoint(None)                                    # valid: same view
    gpu_ctx.set_camera(s.camera)                             # same resolution, new view
    gpu_ctx.render(2, depth=3, seed=1)
    with pytest.raises(eleven.ElevenError):
        gpu_ctx.adjoint(None)


@pytest.mark.gpu
def test_batch_graph_keyed_on_culling(gpu_ctx):
    print("[regress] batch graph: culling vs faithful replays", flush=True)
    sky = scenegen.sky_hdr(64, 32, sun_dir=(0.2, 0.8, 0.1), peak=20.0, sigma_deg=8.0)
    b = eleven.Batch(0)
    for i in range(3):
        b.add(scenegen.sphere_field_arrays(n_spheres=4 + 3 * i, res=(24 + 8 * i, 24), seed=40 + i, hdri=sky), 3,
              40 + i)
    eager_cull = b.render(depth=5).rays
    eager_faithful = b.render(depth=5, cull=False).rays
    assert eager_cull < eager_faithful
    for _ in range(2):
        assert b.render(depth=5, graph=True).rays == eager_cull
        assert b.render(depth=5, graph=True, cull=False).rays == eager_faithful
    b.close()


@pytest.mark.gpu
def test_aov_buffer_grows_with_the_generation():
    print("[regress] wavefront AOV buffer: small AOV render, then a larger one", flush=True)
    small = scenegen.sphere_field_arrays(n_spheres=6, res=(32, 32), seed=5)
    large = scenegen.sphere_field_arrays(n_spheres=6, res=(48, 40), seed=6)
    used = eleven.Context(0)
    try:
        used.upload(small)
        used.render(64, depth=3, seed=2)                      # 65,536 paths: wavefront buffers grow
        used.clear()
        used.render(2, depth=3, seed=2, flags=eleven.FLAG_AOV)   # 2,048 paths: sizes the AOV buffer
        used.upload(large)
        st = used.render(8, depth=3, seed=3, flags=eleven.FLAG_AOV)   # 15,360 paths
        got = [used.read_aov(p) for p in (eleven.PASS_NORMAL, eleven.PASS_TANGENT, eleven.PASS_BITANGENT)]
        got_img = used.read()[0]
    finally:
        used.close()
    fresh = eleven.Context(0)
    try:
        fresh.upload(large)
        fresh.render(8, depth=3, seed=3, flags=eleven.FLAG_AOV)
        ref = [fresh.read_aov(p) for p in (eleven.PASS_NORMAL, eleven.PASS_TANGENT, eleven.PASS_BITANGENT)]
        ref_img = fresh.read()[0]
    finally:
        fresh.close()
    assert st.paths == 48 * 40 * 8
    assert np.array_equal(got_img, ref_img)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)
    assert np.abs(ref[0]).max() > 0.1

"""Host-side checks (CPU only): the C-ABI library loads and exports every
declared symbol, the scene loader keeps the reference loader's semantics, and
the HDRI sampling table reproduces the reference search's distribution."""
import json
import re
from pathlib import Path

import numpy as np
import pytest

import eleven
from eleven import imageio, scenegen
import oracle_py as orc

ROOT = Path(__file__).resolve().parent.parent


# ----------------------------------------------------------------- C ABI --
def test_library_exports_every_declared_symbol():
    hdr = (ROOT / "include" / "eleven.h").read_text()
    decl = set(re.findall(r"\b(el_[a-z_0-9]+)\s*\(", hdr))
    assert len(decl) > 20
    L = eleven.lib()
    missing = [n for n in sorted(decl) if not hasattr(L, n)]
    assert not missing, missing
    assert decl == set(eleven._SIGS), "binding signatures out of sync with the header"


def test_no_device_is_an_error_not_a_crash():
    if eleven.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(eleven.ElevenError) as e:
        eleven.Context(0)
    assert e.value.code == eleven.EL_ERR_HIP


# ---------------------------------------------------------------- loader --
def _write_scene(d: Path, obj: str, mtl: str = "", js: dict = None):
    d.mkdir(parents=True, exist_ok=True)
    (d / "scene.obj").write_text(obj)
    (d / "scene.mtl").write_text(mtl)
    js = js or {"camera": {"xRes": 8, "yRes": 4, "focalLength": 0.05, "focusDistance": 3, "aperture": 2.0,
                           "bokeh": True, "position": {"x": 1, "y": 2, "z": 3}, "rotation": {"x": 10, "y": 20, "z": 30}},
                "hdri": {"color": {"r": 0.1, "g": 0.2, "b": 0.3}},
                "pointLights": [{"position": {"x": 1, "y": 2, "z": 3}, "radiance": {"x": 4, "y": 5, "z": 6}}]}
    (d / "scene.json").write_text(json.dumps(js))


OBJ = """# two objects
v 9 9 9
o first
v 0 0 1
v 1 0 1
v 0 1 1
vt 0 0
vt 1 0
vt 0 1
vn 0 0 2
usemtl red
f 1/1/1 2/2/1 3/3/1
o second
v 0 0 5
v 2 0 5
v 2 2 5
v 0 2 5
usemtl blue
f 4//1 5//1 6//1 7//1
f -4 -3 -2
"""
MTL = """newmtl red
Kd 0.9 0.1 0.1
Ks 0.25 0.3 0.3
Pr 0.4
Pm 1.0
newmtl blue
Kd 0.1 0.1 0.9
Ke 1 2 3
"""


def test_loader_semantics(tmp_path):
    _write_scene(tmp_path / "s", OBJ, MTL)
    s = eleven.SceneData.load(tmp_path / "s")
    v = s.vertices().reshape(-1, 3, 3)
    # lines before the first 'o' are never parsed (ObjLoader.hpp:173-189), so
    # "v 9 9 9" is not vertex 1; z is negated (ObjLoader.hpp:85)
    assert s.tri_count == 1 + 2 + 1
    np.testing.assert_array_equal(v[0], [[0, 0, -1], [1, 0, -1], [0, 1, -1]])
    np.testing.assert_array_equal(v[1], [[0, 0, -5], [2, 0, -5], [2, 2, -5]])   # quad fanned
    np.testing.assert_array_equal(v[2], [[0, 0, -5], [2, 2, -5], [0, 2, -5]])
    np.testing.assert_array_equal(v[3], [[0, 0, -5], [2, 0, -5], [2, 2, -5]])   # negative indices
    n = np.ctypeslib.as_array(s.view.normals, (4, 9))
    np.testing.assert_array_equal(n[0, :3], [0, 0, -1])                          # normalised, z negated
    uv = np.ctypeslib.as_array(s.view.uvs, (4, 6))
    np.testing.assert_array_equal(uv[0], [0, 0, 1, 0, 0, 1])
    oid = np.ctypeslib.as_array(s.view.object_id, (4,))
    assert oid.tolist() == [0, 1, 1, 1]
    mm = np.ctypeslib.as_array(s.view.mesh_material, (2,))
    assert mm.tolist() == [0, 1]
    red, blue = s.materials()
    assert abs(red.albedo.x - 0.9) < 1e-7 and red.specular == 0.25 and abs(red.roughness - 0.4) < 1e-7
    assert red.metallic == 1.0 and blue.roughness == 1.0 and blue.metallic == 0.0     # Material defaults
    assert (blue.emission.x, blue.emission.y, blue.emission.z) == (1, 2, 3)
    c = s.camera
    assert (c.x_res, c.y_res, c.bokeh) == (8, 4, 1)
    assert c.sensor_height == np.float32(np.float32(0.035) * np.float32(4 / 8))
    assert (c.position.x, c.rotation.z) == (1, 30)
    assert s.view.light_count == 1 and s.view.lights[0].radiance.z == 6
    h = s.hdri()
    assert h.shape == (1024, 1024, 3) and np.all(h[..., 1] == np.float32(0.2))       # HDRI(Vector3 color)


def test_loader_textures_gamma_flip_and_hdri_orientation(tmp_path):
    d = tmp_path / "t"
    img = np.zeros((2, 3, 3), np.uint8)
    img[0] = [[255, 0, 0], [0, 255, 0], [0, 0, 255]]       # top row
    img[1] = [[10, 20, 30], [128, 128, 128], [255, 255, 255]]
    d.mkdir()
    scenegen.write_bmp(d / "albedo.bmp", img)
    sky = np.zeros((4, 8, 3), np.float32)
    sky[0] = 2.0                                           # first scanline in the file
    (d / "HDRI").mkdir()
    scenegen.write_hdr(d / "HDRI" / "sky.hdr", sky)
    _write_scene(d, "o a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "newmtl m\nmap_Kd albedo.bmp\n",
                 {"camera": {"xRes": 4, "yRes": 4, "focalLength": 0.035, "focusDistance": 1, "aperture": 2.8,
                             "bokeh": False, "position": {"x": 0, "y": 0, "z": 0},
                             "rotation": {"x": 0, "y": 0, "z": 0}},
                  "hdri": {"name": "sky", "yOffset": 0.25}, "pointLights": []})
    s = eleven.SceneData.load(d)
    t = s.texture(0)
    # flipped on load: data row 0 is the bottom image row (Texture.hpp:49)
    expect = imageio.fast_pow(img[::-1].astype(np.float32) / np.float32(255.0), float(np.float32(2.2)))
    np.testing.assert_array_equal(t, expect.astype(np.float32))
    assert s.materials()[0].albedo_tex == 0
    h = s.hdri()
    # the HDRI is loaded before any texture, so it is not flipped (row 0 = first scanline)
    assert np.all(h[0] == 2.0) and np.all(h[1:] == 0)
    # "yOffset" writes xOffset (SceneLoader.hpp:63-64)
    assert s.view.hdri.x_offset == 0.25 and s.view.hdri.y_offset == 0.0
    cdf, rs = s.hdri_cdf()
    assert rs == 2.0 * 3 * 8 and cdf[8] == 1.0 and cdf[-1] == 1.0


@pytest.mark.parametrize("broken,code", [("missing_obj", eleven.EL_ERR_IO), ("bad_json", eleven.EL_ERR_IO),
                                         ("bad_face", eleven.EL_ERR_IO), ("missing_texture", eleven.EL_ERR_IO),
                                         ("corrupt_hdr", eleven.EL_ERR_IO)])
def test_loader_errors_are_clean(tmp_path, broken, code):
    d = tmp_path / broken
    _write_scene(d, "o a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    if broken == "missing_obj":
        (d / "scene.obj").unlink()
    elif broken == "bad_json":
        (d / "scene.json").write_text("{\"camera\": {\"xRes\": 4,")
    elif broken == "bad_face":
        (d / "scene.obj").write_text("o a\nv 0 0 0\nf 1 2 9\n")
    elif broken == "missing_texture":
        (d / "scene.mtl").write_text("newmtl m\nmap_Kd nothere.bmp\n")
    elif broken == "corrupt_hdr":
        (d / "HDRI").mkdir()
        (d / "HDRI" / "sky.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 4 +X 8\n\x02\x02")
        js = json.loads((d / "scene.json").read_text())
        js["hdri"] = {"name": "sky"}
        (d / "scene.json").write_text(json.dumps(js))
    with pytest.raises(eleven.ElevenError) as e:
        eleven.SceneData.load(d)
    assert e.value.code == code
    assert len(str(e.value)) > 20


def test_sphere_field_workload_size(tmp_path):
    s = eleven.SceneData.load(scenegen.sphere_field(tmp_path / "sphere_field"))
    assert s.tri_count == 61442 and s.view.material_count == 65 and s.view.light_count == 2


# ---------------------------------------------------------- HDRI sampler --
def _maps():
    rng = np.random.default_rng(0)
    out = [np.full((16, 16, 3), 0.5, np.float32)]               # constant: exact CDF ties
    for seed in range(3):
        h, w = int(rng.integers(1, 7)), int(rng.integers(2, 11))
        rgb = (rng.uniform(0, 1, size=(h, w, 3)) * (rng.uniform(size=(h, w, 1)) < 0.7)).astype(np.float32)
        rgb[0, 0] += 0.5
        out.append(rgb)
    sky = scenegen.sky_hdr(64, 32, peak=200.0, sigma_deg=5.0)  # a sun lobe: very uneven mass
    out.append(sky)
    return out


@pytest.mark.parametrize("which", range(5))
def test_hdri_lookup_equals_reference_search(which):
    """The guide-table lookup the kernels run returns the oracle's
    binarySearch(cdf, r) for every 24-bit uniform the renderer can draw
    (exhaustive for the small maps, a dense prefix sweep plus random values for
    the sky), including exact ties r == cdf[j]."""
    rgb = _maps()[which]
    cdf, _ = orc.hdri_cdf(rgb)
    L = rgb.shape[0] * rgb.shape[1]
    if L <= 256:
        ks = np.arange(1 << 24, dtype=np.int64)[::7]
    else:
        ks = np.concatenate([np.arange(1 << 18), np.random.default_rng(1).integers(0, 1 << 24, 1 << 17)])
    rs = (ks.astype(np.float64) / (1 << 24)).astype(np.float32)
    got = eleven.hdri_lookup(rgb, rs)
    want = np.array([orc.hdri_search(cdf, r, L) for r in rs[:: max(1, rs.size // 200000)]])
    assert np.array_equal(got[:: max(1, rs.size // 200000)], want)
    # ties: every CDF value that a 24-bit uniform can equal
    tie = cdf[:L][(cdf[:L] * (1 << 24)) % 1 == 0]
    tie = tie[tie < 1]
    assert np.array_equal(eleven.hdri_lookup(rgb, tie), [orc.hdri_search(cdf, r, L) for r in tie])


def test_hdri_lookup_exhaustive_on_the_workload_sky(tmp_path):
    """The 2048x1024 sky of the heightfield workload (gradient + a 2-degree
    sun lobe of peak 1e3, RGBE-quantised through the .hdr loader, so many
    pixels share values and the CDF has long runs of ties): the renderer's
    guide-table lookup (2^21 cells) returns binarySearch(cdf, r)
    (HDRI.hpp:130-142) for every one of the 2^24 values r the renderer
    draws.  The search's answer is constant for r strictly between two
    consecutive distinct CDF values (every comparison it makes is against a
    CDF value), so the reference answer of every r comes from one search per
    such interval plus one per r equal to a CDF value."""
    import estimator_expectation as ee
    d = tmp_path / "sky_scene"
    d.mkdir()
    w = scenegen.ObjWriter()
    w.add_mesh("t", "m", np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0)], float), np.zeros((3, 2)),
               np.tile([0.0, 0.0, 1.0], (3, 1)), np.array([(0, 1, 2)]))
    w.write(d / "scene.obj")
    scenegen.write_mtl(d / "scene.mtl", [{"name": "m"}])
    (d / "HDRI").mkdir()
    scenegen.write_hdr(d / "HDRI" / "sky.hdr", scenegen.sky_hdr(2048, 1024))
    scenegen.write_json(d / "scene.json", {"xRes": 8, "yRes": 8}, {"name": "sky"}, [])
    scene = eleven.SceneData.load(d)
    rgb = scene.hdri().copy()             # the view lives as long as the scene
    assert rgb.shape == (1024, 2048, 3)
    cdf, _ = orc.hdri_cdf(rgb)
    L = rgb.shape[0] * rgb.shape[1]
    c = np.unique(cdf[:L].astype(np.float64))
    i = np.arange(1 << 24, dtype=np.int64)
    r = i / float(1 << 24)
    j = np.searchsorted(c, r)                          # c[j-1] < r <= c[j]
    tie = (j < c.size) & (c[np.minimum(j, c.size - 1)] == r)
    # one representative grid value per interval (c[j-1], c[j]): the first one inside
    first = np.full(c.size + 1, -1, np.int64)
    inner = ~tie
    jj, pos = np.unique(j[inner], return_index=True)
    first[jj] = i[inner][pos]
    ans_int = np.full(c.size + 1, -1, np.int64)
    ans_int[jj] = ee.binary_search(cdf, (first[jj] / float(1 << 24)).astype(np.float32), L)
    want = ans_int[j]
    want[tie] = ee.binary_search(cdf, r[tie].astype(np.float32), L)
    got = eleven.hdri_lookup(rgb, r.astype(np.float32))
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{bad.size} mismatches, first r = {r[bad[:5]]}"
    assert tie.sum() > 0 and np.unique(want).size > 0.5 * L


def test_hdri_lookup_rejects_out_of_range():
    with pytest.raises(eleven.ElevenError):
        eleven.hdri_lookup(np.ones((2, 2, 3), np.float32), np.array([1.0], np.float32))


# --------------------------------------------------------------- imageio --
def test_fast_pow_pinned_by_hand_computed_values(tmp_path):
    """Ankerl's approximation as the reference's stb_image carries it
    (stb_image.h:127-136): the high word of the double becomes
    (int)(b * (hi - 1072632447) + 1072632447), the low word 0.  Worked by hand:
    fastPow(2, 0.5): hi 0x40000000 -> (int)(0.5 * 1109377 + 1072632447) =
    1073187135 = 0x3FF7893F, i.e. 1 + 0x7893F / 2^20; fastPow(0.5, 2.2): hi
    0x3FE00000 -> 0x3FCDE9CE; fastPow(1, 1/2.2) -> 0x3FEF7E73 (not 1: the
    constant is not 1.0's high word); fastPow(0.25, 1/2.2) -> 0x3FE0F2D0.
    The loader's byte -> value table uses gamma 2.2f (a float,
    stb_image.h:1863): byte 255 -> 0x3FF11D01 (1.0696, not 1), byte 1 ->
    0x3ED7A6BD (0x3ED7A6BE with a double 2.2)."""
    cases = [(2.0, 0.5, "0x1.7893fp+0"), (0.5, 2.2, "0x1.de9cep-3"), (1.0, 1 / 2.2, "0x1.f7e73p-1"),
             (0.25, 1 / 2.2, "0x1.0f2dp-1")]
    for a, b, h in cases:
        assert imageio.fast_pow(np.array([a]), b)[0] == float.fromhex(h), (a, b)
    # the CLI's output mapping (main.cpp:157): (uint8)(fastPow(0.25, 1/2.2) * 255) = 135
    assert imageio.to_ldr(np.full((1, 1, 4), 0.25, np.float32))[0, 0, 0] == 135
    # the loader's table, read back from an sRGB albedo map
    d = tmp_path / "fp"
    d.mkdir()
    grey = np.array([1, 64, 128, 200, 255], np.uint8)
    scenegen.write_bmp(d / "g.bmp", np.repeat(grey[None, :, None], 3, -1))
    _write_scene(d, "o a\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n", "newmtl m\nmap_Kd g.bmp\n")
    t = eleven.SceneData.load(d).texture(0)[0, :, 0].copy()
    want = [float.fromhex(h) for h in ("0x1.7a6bdp-18", "0x1.ad9f1p-5", "0x1.e0d24p-3", "0x1.1edd1p-1",
                                       "0x1.11d01p+0")]
    assert t.tolist() == want


def test_fast_pow_bit_trick_properties():
    a = np.array([1.0, 0.25, 0.5, 0.75])
    assert imageio.fast_pow(a[:1], 1.0)[0] == 1.0                   # low word of 1.0 is zero
    r = imageio.fast_pow(a, 1 / 2.2)
    assert np.all(np.abs(r / a ** (1 / 2.2) - 1) < 0.05)             # published accuracy class
    x = np.linspace(0.01, 1, 200)
    assert np.all(np.diff(imageio.fast_pow(x, 2.2)) >= 0)
    ldr = imageio.to_ldr(np.array([[[0.0, 0.5, 1.0, 2.0]]], np.float32))
    assert ldr[0, 0, 0] == 0 and ldr[0, 0, 2] == ldr[0, 0, 3] and 170 < ldr[0, 0, 1] < 195

"""Tangent frames of the loader against the reference's own MikkTSpace.

The loader computes every corner's tangent and each triangle's sign with
its own implementation of Mikkelsen's method (csrc/tangent_space.cpp), per
mesh object as the reference does (ObjLoader.hpp:167-168,
mikktspaceCallback.hpp:27-99).  oracle/_ref/libmikk_ref.so is the
reference's vendored mikktspace.cpp, compiled where it lies by
oracle/Makefile, driven by oracle/mikk_ref_glue.cpp with the reference's
callbacks (positions, smooth-shading normals, uvs per corner; one sign per
triangle, the last corner's).  Every tangent and sign must be identical bit
for bit, on the workload meshes and on meshes built to reach the method's
special cases: shared-edge fans, a non-manifold edge (four triangles on one
edge), degenerate triangles (two corners on one vertex), a triangle with
zero uv area (joins any group), mirrored uv islands (orientation
reversed), and vertices welded only where position, normal and uv all agree.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import eleven
from eleven import scenegen

REF = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libmikk_ref.so"

pytestmark = pytest.mark.skipif(not REF.exists(), reason="oracle/_ref/libmikk_ref.so is built only where the "
                                                          "reference sources are present (oracle/Makefile)")


def _ref_lib():
    L = C.CDLL(str(REF))
    L.mikk_ref_tangents.argtypes = [C.c_int] + [C.c_void_p] * 5
    L.mikk_ref_tangents.restype = C.c_int
    return L


def _compare(scene):
    v = scene.view
    n = v.tri_count
    get = lambda p, w: np.ctypeslib.as_array(p, (n, w)).copy()
    P, N, UV, T = get(v.vertices, 9), get(v.normals, 9), get(v.uvs, 6), get(v.tangents, 9)
    S = np.ctypeslib.as_array(v.tangent_sign, (n,)).copy()
    O = np.ctypeslib.as_array(v.object_id, (n,)).copy()
    RT, RS = np.zeros_like(T), np.zeros_like(S)
    L = _ref_lib()
    for o in np.unique(O):
        idx = np.flatnonzero(O == o)
        a, b = int(idx[0]), int(idx[-1]) + 1
        assert np.array_equal(idx, np.arange(a, b))           # mesh objects are contiguous
        p, nr, uv = (np.ascontiguousarray(x[a:b]) for x in (P, N, UV))
        t, s = np.zeros((b - a, 9), np.float32), np.zeros(b - a, np.float32)
        assert L.mikk_ref_tangents(b - a, p.ctypes.data, nr.ctypes.data, uv.ctypes.data, t.ctypes.data,
                                   s.ctypes.data) == 0
        RT[a:b], RS[a:b] = t, s
    return T, RT, S, RS


def _soup(path, seed, nv=60, nf=160):
    """Random triangles over shared vertices plus the special cases."""
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(nv, 3))
    VT = np.vstack([rng.uniform(0, 1, size=(nv, 2)), [[0.5, 0.5]]])
    VN = rng.normal(size=(nv, 3))
    faces = [tuple(rng.choice(nv, 3, replace=False)) for _ in range(nf)]
    faces += [(0, 1, 2), (1, 0, 3), (0, 1, 4), (1, 0, 5)]     # four triangles on edge (0, 1)
    faces += [(6, 6, 7), (8, 9, 8)]                          # degenerate
    k = len(VT)
    lines = ["o soup\n"] + [f"v {x} {y} {z}\n" for x, y, z in V] + [f"vt {u} {w}\n" for u, w in VT]
    lines += [f"vn {x} {y} {z}\n" for x, y, z in VN] + ["usemtl m\n"]
    lines += [f"f {a + 1}/{a + 1}/{a + 1} {b + 1}/{b + 1}/{b + 1} {c + 1}/{c + 1}/{c + 1}\n" for a, b, c in faces]
    lines.append(f"f 11/{k}/11 12/{k}/12 13/{k}/13\n")      # zero uv area
    lines += ["o mirror\n", "usemtl m\n"]                    # uvs mirrored: orientation reversed
    lines += [f"f {a + 1}/{c + 1}/{a + 1} {b + 1}/{b + 1}/{b + 1} {c + 1}/{a + 1}/{c + 1}\n" for a, b, c in faces[:40]]
    path.mkdir(parents=True, exist_ok=True)
    (path / "scene.obj").write_text("".join(lines))
    scenegen.write_mtl(path / "scene.mtl", [{"name": "m"}])
    scenegen.write_json(path / "scene.json", {"xRes": 8, "yRes": 8}, {"color": {"r": 1, "g": 1, "b": 1}}, [])
    return path


@pytest.mark.parametrize("kind", ["sphere_field", "heightfield", "soup0", "soup1", "soup2"])
def test_tangents_equal_reference_mikktspace(tmp_path, kind):
    if kind == "sphere_field":
        d = scenegen.sphere_field(tmp_path / kind, n_spheres=8, res=(16, 16), seed=1)
    elif kind == "heightfield":
        d = scenegen.heightfield(tmp_path / kind, cells=(64, 32), extent=(12.8, 6.4), res=(16, 9), n_mats=4,
                                 tex_size=16, sky=(16, 8))
    else:
        d = _soup(tmp_path / kind, int(kind[-1]))
    scene = eleven.SceneData.load(d)
    T, RT, S, RS = _compare(scene)
    assert np.array_equal(T, RT), f"{(T != RT).any(1).sum()} triangles with other tangents"
    assert np.array_equal(S, RS), f"{(S != RS).sum()} triangles with another sign"
    if kind.startswith("soup"):
        assert (S < 0).any() and (S > 0).any()               # both orientations present

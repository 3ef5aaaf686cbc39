"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
always as the checker, never as the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

_lib = None


class or_adjoint_params(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("s0", C.c_uint32), ("depth", C.c_uint32), ("seed", C.c_uint64),
                ("precision", C.c_int), ("shadow_mode", C.c_int), ("threads", C.c_int), ("grad_texture", C.c_int)]


class or_render_params(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("s0", C.c_uint32), ("depth", C.c_uint32), ("seed", C.c_uint64),
                ("precision", C.c_int), ("shadow_mode", C.c_int), ("threads", C.c_int),
                ("x0", C.c_int), ("x1", C.c_int), ("y0", C.c_int), ("y1", C.c_int),
                ("eval_materials", C.c_void_p), ("aov", C.c_void_p)]


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
        L = C.CDLL(str(ORACLE_SO))
        P = C.c_void_p
        L.or_tri_hit.argtypes = [P, P, P, P, P, P]
        L.or_tri_hit.restype = C.c_int
        L.or_closest_brute.argtypes = [P, C.c_int, P, C.c_int, P, P, P, P, C.c_int]
        L.or_refbvh_build.argtypes = [P, C.c_int, C.c_int, C.c_int]
        L.or_refbvh_build.restype = P
        L.or_refbvh_free.argtypes = [P]
        L.or_refbvh_closest.argtypes = [P, P, C.c_int, P, P, C.c_int]
        L.or_refbvh_any.argtypes = [P, P, P, C.c_int, P, C.c_int]
        L.or_refbvh_node_count.argtypes = [P]
        L.or_refbvh_node_count.restype = C.c_int
        L.or_philox.argtypes = [P, P, P]
        L.or_disney_eval.argtypes = [P, P, P, C.c_int, P]
        L.or_disney_pdf.argtypes = [P, P, P, C.c_int]
        L.or_disney_pdf.restype = C.c_double
        L.or_disney_sample.argtypes = [P, P, C.c_double, C.c_double, C.c_double, C.c_int, P]
        L.or_cosine_sample_hemisphere.argtypes = [C.c_double, C.c_double, P]
        L.or_importance_sample_ggx.argtypes = [C.c_double, C.c_double, C.c_double, P]
        L.or_uniform_circle.argtypes = [C.c_double, C.c_double, C.c_double, P, P]
        L.or_hdri_cdf.argtypes = [P, C.c_int, C.c_int, P, P]
        L.or_hdri_search.argtypes = [P, C.c_float, C.c_int]
        L.or_hdri_search.restype = C.c_int
        L.or_hdri_pdf.argtypes = [P, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int]
        L.or_hdri_pdf.restype = C.c_float
        L.or_camera_ray.argtypes = [P, C.c_int, C.c_int, P, P, P]
        L.or_render.argtypes = [P, P, C.POINTER(or_render_params), P, P, P, C.POINTER(C.c_uint64),
                                C.POINTER(C.c_uint64)]
        L.or_render.restype = C.c_int
        L.or_disney_eval_grad.argtypes = [P, P, P, P]
        L.or_adjoint.argtypes = [P, P, C.POINTER(or_adjoint_params), P, P, P, P]
        L.or_adjoint.restype = C.c_int
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def tri_hit(v9, o, d):
    # keep every temporary array referenced for the duration of the call
    a, b, c = _f32(v9), _f32(o), _f32(d)
    t, u, v = C.c_float(), C.c_float(), C.c_float()
    h = lib().or_tri_hit(a.ctypes.data, b.ctypes.data, c.ctypes.data, C.byref(t), C.byref(u), C.byref(v))
    return bool(h), t.value, u.value, v.value


def closest_brute(verts, rays, threads=0):
    v = _f32(verts).reshape(-1, 9)
    r = _f32(rays).reshape(-1, 6)
    n = r.shape[0]
    tri = np.zeros(n, np.int32)
    t = np.zeros(n, np.float32)
    u = np.zeros(n, np.float32)
    w = np.zeros(n, np.float32)
    lib().or_closest_brute(v.ctypes.data, v.shape[0], r.ctypes.data, n, tri.ctypes.data, t.ctypes.data,
                           u.ctypes.data, w.ctypes.data, threads)
    return tri, t, u, w


class RefBVH:
    """Reference-faithful depth-18, 14-bin SAH binary BVH (BVH.hpp)."""

    def __init__(self, verts, depth=18, bins=14):
        self.v = _f32(verts).reshape(-1, 9)
        self.h = lib().or_refbvh_build(self.v.ctypes.data, self.v.shape[0], depth, bins)
        if not self.h:
            raise ValueError("bad BVH parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_refbvh_free(self.h)
            self.h = None

    def closest(self, rays, threads=0):
        r = _f32(rays).reshape(-1, 6)
        tri = np.zeros(r.shape[0], np.int32)
        t = np.zeros(r.shape[0], np.float32)
        lib().or_refbvh_closest(self.h, r.ctypes.data, r.shape[0], tri.ctypes.data, t.ctypes.data, threads)
        return tri, t

    def any(self, rays, tmax, threads=0):
        r = _f32(rays).reshape(-1, 6)
        tm = _f32(tmax)
        occ = np.zeros(r.shape[0], np.uint8)
        lib().or_refbvh_any(self.h, r.ctypes.data, tm.ctypes.data, r.shape[0], occ.ctypes.data, threads)
        return occ.astype(bool)


def philox(counter, key):
    c = np.ascontiguousarray(counter, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().or_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


HD_FIELDS = ["metallic", "roughness", "clearcoatGloss", "clearcoat", "anisotropic", "eta", "transmission",
             "specular", "specularTint", "sheenTint", "subsurface", "sheen"]


def hitdata(normal=(0, 0, 1), albedo=(0.5, 0.5, 0.5), emission=(0, 0, 0), **kw):
    """HitData as 21 doubles: 12 scalars (HD_FIELDS order), emission, albedo, normal."""
    d = {"metallic": 0.0, "roughness": 0.5, "clearcoatGloss": 0.0, "clearcoat": 0.0, "anisotropic": 0.0,
         "eta": 0.0, "transmission": 0.0, "specular": 0.5, "specularTint": 0.0, "sheenTint": 0.0,
         "subsurface": 0.0, "sheen": 0.0}
    d.update(kw)
    return np.array([d[k] for k in HD_FIELDS] + list(emission) + list(albedo) + list(normal), np.float64)


def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def disney_eval(hd, direction, L, precision=1):
    out = np.zeros(3)
    h, d, l = _f64(hd), _f64(direction), _f64(L)
    lib().or_disney_eval(h.ctypes.data, d.ctypes.data, l.ctypes.data, precision, out.ctypes.data)
    return out


def disney_eval_grad(hd, direction, L):
    """FP64 d f / d (albedo r, g, b, roughness, metallic, specular): [6][3]."""
    out = np.zeros(18)
    h, d, l = _f64(hd), _f64(direction), _f64(L)
    lib().or_disney_eval_grad(h.ctypes.data, d.ctypes.data, l.ctypes.data, out.ctypes.data)
    return out.reshape(6, 3)


def disney_pdf(hd, direction, L, precision=1):
    h, d, l = _f64(hd), _f64(direction), _f64(L)
    return lib().or_disney_pdf(h.ctypes.data, d.ctypes.data, l.ctypes.data, precision)


def disney_sample(hd, direction, r1, r2, r3, precision=1):
    out = np.zeros(3)
    h, d = _f64(hd), _f64(direction)
    lib().or_disney_sample(h.ctypes.data, d.ctypes.data, r1, r2, r3, precision, out.ctypes.data)
    return out


def cosine_sample_hemisphere(u1, u2):
    out = np.zeros(3)
    lib().or_cosine_sample_hemisphere(u1, u2, out.ctypes.data)
    return out


def importance_sample_ggx(rgh, r1, r2):
    out = np.zeros(3)
    lib().or_importance_sample_ggx(rgh, r1, r2, out.ctypes.data)
    return out


def uniform_circle(u1, u2, u3):
    x, y = C.c_double(), C.c_double()
    lib().or_uniform_circle(u1, u2, u3, C.byref(x), C.byref(y))
    return x.value, y.value


def hdri_cdf(rgb):
    a = _f32(rgb)
    h, w = a.shape[:2]
    cdf = np.zeros(w * h + 1, np.float32)
    rs = C.c_float()
    lib().or_hdri_cdf(a.ctypes.data, w, h, cdf.ctypes.data, C.byref(rs))
    return cdf, rs.value


def hdri_search(cdf, r, length):
    c = _f32(cdf)
    return lib().or_hdri_search(c.ctypes.data, float(r), length)


def hdri_pdf(rgb, x_offset, radiance_sum, x, y):
    a = _f32(rgb)
    h, w = a.shape[:2]
    return lib().or_hdri_pdf(a.ctypes.data, w, h, x_offset, radiance_sum, x, y)


def camera_ray(cam, x, y, r5):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    r = _f32(r5)
    lib().or_camera_ray(C.byref(cam), x, y, r.ctypes.data, o.ctypes.data, d.ctypes.data)
    return o, d


def render(scene, refbvh, spp, depth=5, seed=0, s0=0, precision=0, shadow_mode=1, threads=0, window=None,
           eval_materials=None, aov=None):
    """Oracle render of pixels `window` = (x0, x1, y0, y1) (kernel x, y).

    eval_materials: optional ctypes array of el_material whose BRDF values
    are used while sampling, pdfs and MIS weights keep the scene's.
    aov: optional float64 array [H, W, 9] that receives the sums of the first
    hit's normal, tangent and bitangent over the accepted samples.
    Returns (sum[H,W,3], sumsq[H,W,3], count[H,W], rays, vertices) in output
    layout (row 0 = top)."""
    cam = scene.view.camera
    W, H = cam.x_res, cam.y_res
    x0, x1, y0, y1 = window or (0, W, 0, H)
    p = or_render_params()
    p.spp, p.s0, p.depth, p.seed = spp, s0, depth, seed
    p.precision, p.shadow_mode, p.threads = precision, shadow_mode, threads
    p.x0, p.x1, p.y0, p.y1 = x0, x1, y0, y1
    p.eval_materials = C.cast(eval_materials, C.c_void_p) if eval_materials is not None else None
    if aov is not None:
        assert aov.dtype == np.float64 and aov.shape == (H, W, 9) and aov.flags.c_contiguous
        p.aov = aov.ctypes.data
    s = np.zeros((H, W, 3))
    ss = np.zeros((H, W, 3))
    cnt = np.zeros((H, W), np.uint32)
    rays, verts = C.c_uint64(), C.c_uint64()
    rc = lib().or_render(C.byref(scene.view), refbvh.h, C.byref(p), s.ctypes.data, ss.ctypes.data,
                         cnt.ctypes.data, C.byref(rays), C.byref(verts))
    if rc != 0:
        raise ValueError("or_render rejected its arguments")
    return s, ss, cnt, rays.value, verts.value


def adjoint(scene, refbvh, dimage, spp, depth=5, seed=0, s0=0, precision=1, shadow_mode=1, threads=0,
            grad_texture=-1):
    """Path-replay adjoint on the CPU: returns (dmat [M][6], dtex [h][w][3] or None, image [H][W][3])
    for the loss sum(dimage * I); dimage and image in the reference's idx order
    (row 0 = top, as el_read_passes)."""
    v = scene.view
    H, W = v.camera.y_res, v.camera.x_res
    di = np.ascontiguousarray(dimage, np.float32).reshape(H * W * 3)
    dmat = np.zeros((v.material_count, 6))
    dtex = None
    if grad_texture >= 0:
        t = v.textures[grad_texture]
        dtex = np.zeros((t.height, t.width, 3))
    img = np.zeros((H, W, 3))
    p = or_adjoint_params(spp, s0, depth, seed, precision, shadow_mode, threads, grad_texture)
    rc = lib().or_adjoint(C.byref(v), refbvh.h, C.byref(p), di.ctypes.data, dmat.ctypes.data,
                          dtex.ctypes.data if dtex is not None else None, img.ctypes.data)
    if rc != 0:
        raise ValueError("or_adjoint rejected its arguments")
    return dmat, dtex, img

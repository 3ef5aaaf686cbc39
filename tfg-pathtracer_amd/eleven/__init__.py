"""Python binding of the MI355X path-tracing core (ctypes over include/eleven.h).

The product path is ``libeleven.so`` (HIP kernels + host loader/BVH).  This
module only marshals numpy arrays into the C structs and fails loudly when the
library is missing: there is no CPU fallback.

Reference surface mirrored: ``loadScene`` (SceneLoader.hpp:7), ``renderSetup``
/ ``renderCuda`` / ``getBuffers`` / ``getSamples`` (kernel.h:78-84).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent          # tfg-pathtracer_amd/
LIB_PATH = PKG_DIR / "libeleven.so"

EL_OK, EL_ERR_ARG, EL_ERR_IO, EL_ERR_HIP, EL_ERR_STATE, EL_ERR_RANGE, EL_ERR_NOMEM = 0, -1, -2, -3, -4, -5, -6
MODE_PLAIN, MODE_WAVEFRONT = 0, 1
FLAG_AOV, FLAG_COUNT_TRAVERSAL, FLAG_NO_CULL, FLAG_RECORD_PATHS, FLAG_SERIAL_SHADOWS = 1, 2, 4, 8, 16
FLAG_KERNEL_TIMES = 32   # per-launch event timing (el_stats extend_ms, shade_ms, shadow_ms)
ADJ_ACCUMULATE, ADJ_DEVICE_RESIDUAL = 1, 2
PASS_BEAUTY, PASS_DENOISE, PASS_NORMAL, PASS_TANGENT, PASS_BITANGENT = 0, 1, 2, 3, 4


class ElevenError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"eleven error {code}: {msg}")
        self.code = code


class float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class el_material(C.Structure):
    _fields_ = [("albedo_tex", C.c_int32), ("emission_tex", C.c_int32), ("roughness_tex", C.c_int32),
                ("metallic_tex", C.c_int32), ("normal_tex", C.c_int32), ("opacity_tex", C.c_int32),
                ("albedo", float3), ("emission", float3), ("opacity", float3),
                ("roughness", C.c_float), ("metallic", C.c_float), ("clearcoat_gloss", C.c_float),
                ("clearcoat", C.c_float), ("anisotropic", C.c_float), ("eta", C.c_float),
                ("transmission", C.c_float), ("specular", C.c_float), ("specular_tint", C.c_float),
                ("sheen_tint", C.c_float), ("subsurface", C.c_float), ("sheen", C.c_float)]


class el_texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("data", C.POINTER(C.c_float)),
                ("x_tile", C.c_float), ("y_tile", C.c_float), ("x_offset", C.c_float),
                ("y_offset", C.c_float), ("filter", C.c_int32)]


class el_point_light(C.Structure):
    _fields_ = [("position", float3), ("radiance", float3)]


class el_camera(C.Structure):
    _fields_ = [("x_res", C.c_uint32), ("y_res", C.c_uint32), ("focal_length", C.c_float),
                ("sensor_width", C.c_float), ("sensor_height", C.c_float), ("aperture", C.c_float),
                ("focus_distance", C.c_float), ("rotation", float3), ("position", float3),
                ("bokeh", C.c_int32)]


class el_scene_view(C.Structure):
    _fields_ = [("tri_count", C.c_uint32),
                ("vertices", C.POINTER(C.c_float)), ("uvs", C.POINTER(C.c_float)),
                ("normals", C.POINTER(C.c_float)), ("tangents", C.POINTER(C.c_float)),
                ("tangent_sign", C.POINTER(C.c_float)), ("object_id", C.POINTER(C.c_int32)),
                ("mesh_count", C.c_uint32), ("mesh_material", C.POINTER(C.c_int32)),
                ("material_count", C.c_uint32), ("materials", C.POINTER(el_material)),
                ("texture_count", C.c_uint32), ("textures", C.POINTER(el_texture)),
                ("hdri", el_texture),
                ("light_count", C.c_uint32), ("lights", C.POINTER(el_point_light)),
                ("camera", el_camera)]


class el_render_params(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("seed", C.c_uint64), ("mode", C.c_int32),
                ("block", C.c_uint32), ("queue_cap", C.c_uint32), ("bvh_width", C.c_uint32),
                ("flags", C.c_uint32), ("refill", C.c_uint32), ("rng_scene", C.c_uint32)]


class el_batch_params(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("block", C.c_uint32), ("queue_cap", C.c_uint32),
                ("bvh_width", C.c_uint32), ("refill", C.c_uint32), ("flags", C.c_uint32)]


class el_stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("vertices", C.c_uint64), ("paths", C.c_uint64),
                ("render_ms", C.c_double), ("extend_ms", C.c_double), ("shadow_ms", C.c_double),
                ("shade_ms", C.c_double), ("extend_launches", C.c_uint64), ("extend_rays", C.c_uint64),
                ("node_fetches", C.c_uint64), ("tri_tests", C.c_uint64),
                ("node_bytes", C.c_uint64), ("tri_bytes", C.c_uint64), ("bvh_width", C.c_uint32),
                ("stack_bound", C.c_uint32), ("node_stride", C.c_uint32), ("tri_stride", C.c_uint32),
                ("adjoint_ms", C.c_double), ("adjoint_rays", C.c_uint64), ("texels_8bit", C.c_uint64),
                ("adjoint_bytes", C.c_uint64), ("record_bytes", C.c_uint64)]


class el_adjoint_params(C.Structure):
    _fields_ = [("grad_texture", C.c_int32), ("block", C.c_uint32), ("flags", C.c_uint32)]


# Every symbol declared in include/eleven.h, with its ctypes signature.
_P = C.c_void_p
_SIGS = {
    "el_scene_load": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "el_scene_view_get": (C.c_int, [_P, C.POINTER(el_scene_view)]),
    "el_scene_hdri_cdf": (C.c_int, [_P, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_float)]),
    "el_scene_free": (None, [_P]),
    "el_hdri_lookup": (C.c_int, [C.POINTER(el_texture), _P, C.c_uint32, _P]),
    "el_scene_last_error": (C.c_char_p, []),
    "el_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "el_destroy": (None, [_P]),
    "el_last_error": (C.c_char_p, [_P]),
    "el_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "el_upload_scene": (C.c_int, [_P, C.POINTER(el_scene_view)]),
    "el_render": (C.c_int, [_P, C.POINTER(el_render_params)]),
    "el_clear": (C.c_int, [_P]),
    "el_read_passes": (C.c_int, [_P, _P, _P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "el_read_aov": (C.c_int, [_P, C.c_int, _P]),
    "el_get_stats": (C.c_int, [_P, C.POINTER(el_stats)]),
    "el_set_materials": (C.c_int, [_P, C.POINTER(el_material), C.c_uint32]),
    "el_set_texture": (C.c_int, [_P, C.c_int32, _P]),
    "el_set_camera": (C.c_int, [_P, _P]),
    "el_adjoint": (C.c_int, [_P, C.POINTER(el_adjoint_params), _P, _P, _P, C.c_int]),
    "el_residual_l2": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "el_batch_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "el_batch_destroy": (None, [_P]),
    "el_batch_last_error": (C.c_char_p, [_P]),
    "el_batch_log": (C.c_char_p, [_P]),
    "el_batch_add": (C.c_int, [_P, C.POINTER(el_scene_view), C.c_uint32, C.c_uint64, C.POINTER(C.c_int32)]),
    "el_batch_render": (C.c_int, [_P, C.POINTER(el_batch_params)]),
    "el_batch_tune": (C.c_int, [_P, C.POINTER(el_batch_params), C.c_uint32]),
    "el_batch_read": (C.c_int, [_P, C.c_int32, _P, _P]),
    "el_batch_stats": (C.c_int, [_P, C.POINTER(el_stats)]),
    "el_trace_closest": (C.c_int, [_P, _P, C.c_uint32, _P, _P, _P, C.c_int]),
    "el_trace_any": (C.c_int, [_P, _P, _P, C.c_uint32, _P, C.c_int]),
    "el_trace_count": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int]),
    "el_trace_bench": (C.c_int, [_P, _P, C.c_uint32, _P, _P, C.c_int, C.POINTER(C.c_double)]),
    "el_trace_closest_noalloc": (C.c_int, [_P, _P, C.c_uint32, _P, _P, _P]),
    "el_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "el_device_free": (C.c_int, [_P, _P]),
    "el_memcpy_htod": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "el_memcpy_dtoh": (C.c_int, [_P, _P, _P, C.c_size_t]),
}

_lib = None


def lib():
    """Load libeleven.so (build it with `make -C tfg-pathtracer_amd`).  The
    environment variable ELEVEN_LIBRARY names another build of the same
    library (e.g. one compiled with different kernel tuning macros)."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("ELEVEN_LIBRARY") or LIB_PATH)
        if not path.exists():
            raise ElevenError(EL_ERR_STATE, f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, ctype=C.c_float):
    return a.ctypes.data_as(C.POINTER(ctype))


def hdri_lookup(rgb: np.ndarray, r: np.ndarray) -> np.ndarray:
    """Environment pixel index the renderer draws for each uniform r (host evaluation)."""
    t = make_texture(rgb)
    rr = np.ascontiguousarray(r, np.float32)
    out = np.zeros(rr.size, np.int32)
    rc = lib().el_hdri_lookup(C.byref(t), rr.ctypes.data, rr.size, out.ctypes.data)
    if rc != EL_OK:
        raise ElevenError(rc, "el_hdri_lookup failed")
    return out


def device_count() -> int:
    n = C.c_int(0)
    lib().el_device_count(C.byref(n))
    return n.value


# --------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------
class SceneData:
    """A scene as flat numpy arrays plus the el_scene_view pointing at them."""

    def __init__(self):
        self._keep = []
        self.view = el_scene_view()
        self._handle = None

    # ---- construction -----------------------------------------------------
    @classmethod
    def load(cls, directory: str | os.PathLike) -> "SceneData":
        """loadScene(dir) through the C loader (SceneLoader.hpp:7)."""
        s = cls()
        h = C.c_void_p()
        rc = lib().el_scene_load(str(directory).encode(), C.byref(h))
        if rc != EL_OK:
            raise ElevenError(rc, lib().el_scene_last_error().decode())
        s._handle = h
        lib().el_scene_view_get(h, C.byref(s.view))
        return s

    @classmethod
    def from_arrays(cls, vertices, camera: dict, *, normals=None, uvs=None, tangents=None, tangent_sign=None,
                    object_id=None, mesh_material=None, materials=None, textures=(), hdri=None, lights=()):
        s = cls()
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 9)
        n = v.shape[0]
        nrm = np.ascontiguousarray(normals if normals is not None else np.zeros((n, 9)), np.float32).reshape(-1, 9)
        uv = np.ascontiguousarray(uvs if uvs is not None else np.zeros((n, 6)), np.float32).reshape(-1, 6)
        tg = np.ascontiguousarray(tangents if tangents is not None else np.zeros((n, 9)), np.float32).reshape(-1, 9)
        sg = np.ascontiguousarray(tangent_sign if tangent_sign is not None else np.ones(n), np.float32)
        oid = np.ascontiguousarray(object_id if object_id is not None else np.zeros(n), np.int32)
        mm = np.ascontiguousarray(mesh_material if mesh_material is not None else [0], np.int32)
        mats = materials if materials is not None else [make_material()]
        marr = (el_material * len(mats))(*mats)
        texs = [t if isinstance(t, el_texture) else make_texture(t) for t in textures]   # arrays or make_texture()s
        tarr = (el_texture * max(1, len(texs)))(*texs)
        if hdri is None:
            hdri = np.full((1, 1, 3), 0.0, np.float32)
        htex = make_texture(hdri)
        larr = (el_point_light * max(1, len(lights)))(*[
            el_point_light(float3(*l[0]), float3(*l[1])) for l in lights])
        s._keep += [v, nrm, uv, tg, sg, oid, mm, marr, tarr, htex, larr, texs]
        vw = s.view
        vw.tri_count = n
        vw.vertices, vw.normals, vw.uvs, vw.tangents = _ptr(v), _ptr(nrm), _ptr(uv), _ptr(tg)
        vw.tangent_sign = _ptr(sg)
        vw.object_id = _ptr(oid, C.c_int32)
        vw.mesh_count, vw.mesh_material = len(mm), _ptr(mm, C.c_int32)
        vw.material_count, vw.materials = len(mats), C.cast(marr, C.POINTER(el_material))
        vw.texture_count, vw.textures = len(texs), C.cast(tarr, C.POINTER(el_texture))
        vw.hdri = htex
        vw.light_count, vw.lights = len(lights), C.cast(larr, C.POINTER(el_point_light))
        vw.camera = make_camera(**camera)
        return s

    def __del__(self):
        if self._handle is not None and _lib is not None:
            _lib.el_scene_free(self._handle)
            self._handle = None

    # ---- accessors (numpy views of the flat arrays) -------------------------
    # The views point into memory this SceneData owns (the loader's or the
    # arrays from_arrays keeps); each view holds a reference to the SceneData,
    # so the memory lives as long as any view of it.
    def _view(self, ptr, shape, ctype=C.c_float, dtype=np.float32):
        n = int(np.prod(shape))
        if n == 0 or not ptr:
            return np.zeros(shape, dtype)
        buf = (ctype * n).from_address(C.addressof(ptr.contents))
        buf._owner = self
        return np.frombuffer(buf, dtype=dtype).reshape(shape)

    @property
    def tri_count(self) -> int:
        return self.view.tri_count

    def vertices(self) -> np.ndarray:
        return self._view(self.view.vertices, (self.view.tri_count, 9))

    def materials(self) -> list:
        return [self.view.materials[i] for i in range(self.view.material_count)]

    def texture(self, i: int) -> np.ndarray:
        t = self.view.textures[i]
        return self._view(t.data, (t.height, t.width, 3))

    def hdri(self) -> np.ndarray:
        t = self.view.hdri
        return self._view(t.data, (t.height, t.width, 3))

    def hdri_cdf(self):
        """(cdf[w*h+1], radianceSum) as built by HDRI::generateCDF; loaded scenes only."""
        if self._handle is None:
            raise ElevenError(EL_ERR_STATE, "hdri_cdf is available for loaded scenes")
        p = C.POINTER(C.c_float)()
        rs = C.c_float()
        lib().el_scene_hdri_cdf(self._handle, C.byref(p), C.byref(rs))
        t = self.view.hdri
        return np.ctypeslib.as_array(p, (t.width * t.height + 1,)).copy(), rs.value

    @property
    def camera(self) -> el_camera:
        return self.view.camera


def make_material(albedo=(0.5, 0.5, 0.5), roughness=1.0, metallic=0.0, specular=0.5, emission=(0, 0, 0),
                  albedo_tex=-1, **kw) -> el_material:
    """Material.hpp defaults (Material.hpp:14-36) with overrides."""
    m = el_material()
    m.albedo_tex, m.emission_tex, m.roughness_tex = albedo_tex, -1, -1
    m.metallic_tex, m.normal_tex, m.opacity_tex = -1, -1, -1
    m.albedo, m.emission, m.opacity = float3(*albedo), float3(*emission), float3(1, 1, 1)
    m.roughness, m.metallic, m.specular = roughness, metallic, specular
    for k in ("clearcoat_gloss", "clearcoat", "anisotropic", "eta", "transmission", "specular_tint",
              "sheen_tint", "subsurface", "sheen"):
        setattr(m, k, kw.pop(k, 0.0))
    for k, val in kw.items():
        setattr(m, k, val)
    return m


def make_texture(rgb, x_tile=1.0, y_tile=1.0, x_offset=0.0, y_offset=0.0, filt=0) -> el_texture:
    a = np.ascontiguousarray(rgb, np.float32)
    t = el_texture()
    t.height, t.width = a.shape[0], a.shape[1]
    t.data = _ptr(a)
    t.x_tile, t.y_tile, t.x_offset, t.y_offset, t.filter = x_tile, y_tile, x_offset, y_offset, filt
    t._keep = a
    return t


def make_camera(x_res, y_res, position=(0, 0, 0), rotation=(0, 0, 0), focal_length=0.035, sensor_width=0.035,
                aperture=2.8, focus_distance=1e6, bokeh=False) -> el_camera:
    """Camera(xRes, yRes) semantics (Camera.hpp:25-33): sensorHeight follows the aspect."""
    c = el_camera()
    c.x_res, c.y_res = x_res, y_res
    c.focal_length, c.sensor_width = focal_length, sensor_width
    c.sensor_height = np.float32(np.float32(sensor_width) * (np.float32(y_res) / np.float32(x_res)))
    c.aperture, c.focus_distance = aperture, focus_distance
    c.rotation, c.position = float3(*rotation), float3(*position)
    c.bokeh = 1 if bokeh else 0
    return c


# --------------------------------------------------------------------------
# device context
# --------------------------------------------------------------------------
class Context:
    """One device, one stream (el_create).  Methods raise ElevenError on failure."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().el_create(C.byref(self._h), device)
        if rc != EL_OK:
            raise ElevenError(rc, "el_create failed (no HIP device?)")
        self.scene = None
        self.camera = None   # the camera rendered from (the scene's, or set_camera's)

    def _check(self, rc):
        if rc != EL_OK:
            raise ElevenError(rc, lib().el_last_error(self._h).decode())

    def close(self):
        if self._h:
            lib().el_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: SceneData):
        self._check(lib().el_upload_scene(self._h, C.byref(scene.view)))
        self.scene = scene
        self.camera = scene.camera

    def render(self, spp, depth=5, seed=0, mode=MODE_WAVEFRONT, block=0, queue_cap=0, bvh_width=0, flags=0,
               refill=0, rng_scene=0):
        p = el_render_params(spp, depth, seed, mode, block, queue_cap, bvh_width, flags, refill, rng_scene)
        self._check(lib().el_render(self._h, C.byref(p)))
        return self.stats()

    def clear(self):
        self._check(lib().el_clear(self._h))

    def read(self):
        cam = self.camera
        rgba = np.zeros((cam.y_res, cam.x_res, 4), np.float32)
        samples = np.zeros((cam.y_res, cam.x_res), np.uint32)
        rays, verts = C.c_uint64(), C.c_uint64()
        self._check(lib().el_read_passes(self._h, rgba.ctypes.data, samples.ctypes.data, C.byref(rays),
                                         C.byref(verts)))
        return rgba, samples, rays.value, verts.value

    def read_aov(self, pass_id):
        cam = self.camera
        rgba = np.zeros((cam.y_res, cam.x_res, 4), np.float32)
        self._check(lib().el_read_aov(self._h, pass_id, rgba.ctypes.data))
        return rgba

    def stats(self) -> el_stats:
        s = el_stats()
        self._check(lib().el_get_stats(self._h, C.byref(s)))
        return s

    def set_materials(self, mats):
        """mats: el_material structs (a list, or a ctypes array passed as is)."""
        arr = mats if isinstance(mats, C.Array) else (el_material * len(mats))(*mats)
        self._check(lib().el_set_materials(self._h, arr, len(mats)))

    def residual_l2(self, target) -> float:
        """el_residual_l2: keeps 2 (I - target) on the device for an
        adjoint(None, ...) and returns sum((I - target)^2) (I = the mean image
        of read(), target [H][W][3])."""
        cam = self.camera
        t = np.ascontiguousarray(target, np.float32).reshape(cam.y_res, cam.x_res, 3)
        loss = C.c_double()
        self._check(lib().el_residual_l2(self._h, t.ctypes.data, C.byref(loss)))
        return loss.value

    def adjoint(self, dimage, grad_texture=-1, block=0, accumulate=False, read=True):
        """Path-replay adjoint of the accumulated image: gradients of
        sum(dimage * I) (dimage [H][W][3], the layout of read()) with respect to
        every material's (albedo r, g, b, roughness, metallic, specular) and the
        texels of texture `grad_texture`.  Returns (dmat [M][6], dtex [h][w][3]
        or None).  With `accumulate` the gradients add to those of the previous
        calls (EL_ADJ_ACCUMULATE) and the running sums are returned; with
        read=False nothing is copied out and (None, None) is returned (the
        sums stay on the device for following accumulating calls).  With
        dimage None the residual of the last residual_l2 is used
        (EL_ADJ_DEVICE_RESIDUAL)."""
        cam = self.camera
        di = None if dimage is None else np.ascontiguousarray(dimage, np.float32).reshape(cam.y_res, cam.x_res, 3)
        m = len(self.scene.materials())
        dmat = dtex = None
        if read:
            dmat = np.zeros((m, 6), np.float32)
            if grad_texture >= 0:
                t = self.scene.view.textures[grad_texture]
                dtex = np.zeros((t.height, t.width, 3), np.float32)
        flags = (ADJ_ACCUMULATE if accumulate else 0) | (ADJ_DEVICE_RESIDUAL if di is None else 0)
        p = el_adjoint_params(grad_texture, block, flags)
        self._check(lib().el_adjoint(self._h, C.byref(p), di.ctypes.data if di is not None else None,
                                     dmat.ctypes.data if dmat is not None else None,
                                     dtex.ctypes.data if dtex is not None else None, 0))
        return dmat, dtex

    def set_texture(self, tex_id, rgb):
        a = np.ascontiguousarray(rgb, np.float32)
        self._check(lib().el_set_texture(self._h, tex_id, a.ctypes.data))

    def set_camera(self, camera: "el_camera"):
        """Replace the camera (resolution may change) and clear the accumulation."""
        self._check(lib().el_set_camera(self._h, C.byref(camera)))
        self.camera = camera

    def trace_closest(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = r.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        uv = np.zeros((n, 2), np.float32)
        self._check(lib().el_trace_closest(self._h, r.ctypes.data, n, tri.ctypes.data, t.ctypes.data,
                                           uv.ctypes.data, 0))
        return tri, t, uv

    def trace_any(self, rays, tmax):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        tm = np.ascontiguousarray(tmax, np.float32)
        occ = np.zeros(r.shape[0], np.uint8)
        self._check(lib().el_trace_any(self._h, r.ctypes.data, tm.ctypes.data, r.shape[0], occ.ctypes.data, 0))
        return occ.astype(bool)

    def trace_count(self, rays):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        nodes, tris = C.c_uint64(), C.c_uint64()
        self._check(lib().el_trace_count(self._h, r.ctypes.data, r.shape[0], C.byref(nodes), C.byref(tris), 0))
        return nodes.value, tris.value

    # device buffers for resident-input measurements
    def alloc(self, nbytes):
        p = C.c_void_p()
        self._check(lib().el_device_alloc(self._h, nbytes, C.byref(p)))
        return p

    def free(self, p):
        self._check(lib().el_device_free(self._h, p))

    def htod(self, dst, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        self._check(lib().el_memcpy_htod(self._h, dst, a.ctypes.data, a.nbytes))

    def dtoh(self, arr: np.ndarray, src):
        self._check(lib().el_memcpy_dtoh(self._h, arr.ctypes.data, src, arr.nbytes))

    def trace_closest_noalloc(self, d_rays, n, d_tri, d_t, d_uv=None):
        """el_trace_closest_noalloc: device pointers in and out, no allocation."""
        self._check(lib().el_trace_closest_noalloc(self._h, d_rays, n, d_tri, d_t, d_uv))

    def trace_bench(self, d_rays, n, d_tri, d_t, iters):
        ms = C.c_double()
        self._check(lib().el_trace_bench(self._h, d_rays, n, d_tri, d_t, iters, C.byref(ms)))
        return ms.value


BATCH_GRAPH, BATCH_NO_CULL, BATCH_SERIAL_SHADOWS = 1, 4, 8


class Batch:
    """Many scenes through one wavefront loop (el_batch_*)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        rc = lib().el_batch_create(C.byref(h), device)
        if rc != EL_OK:
            raise ElevenError(rc, "el_batch_create failed (no HIP device?)")
        self._h = h
        self.scenes = []

    def _check(self, rc):
        if rc != EL_OK:
            raise ElevenError(rc, lib().el_batch_last_error(self._h).decode())

    def add(self, scene: SceneData, spp: int, seed: int = 0) -> int:
        idx = C.c_int32()
        self._check(lib().el_batch_add(self._h, C.byref(scene.view), spp, seed, C.byref(idx)))
        self.scenes.append(scene)
        return idx.value

    def render(self, depth=5, block=0, queue_cap=0, bvh_width=0, refill=0, graph=False, cull=True,
               serial_shadows=False) -> el_stats:
        flags = (BATCH_GRAPH if graph else 0) | (0 if cull else BATCH_NO_CULL) | \
            (BATCH_SERIAL_SHADOWS if serial_shadows else 0)
        p = el_batch_params(depth, block, queue_cap, bvh_width, refill, flags)
        self._check(lib().el_batch_render(self._h, C.byref(p)))
        return self.stats()

    def tune(self, depth=5, trial_spp=2, graph=False) -> dict:
        p = el_batch_params(depth, 0, 0, 0, 0, BATCH_GRAPH if graph else 0)
        self._check(lib().el_batch_tune(self._h, C.byref(p), trial_spp))
        return {"block": p.block, "queue_cap": p.queue_cap, "bvh_width": p.bvh_width, "refill": p.refill}

    def log(self) -> str:
        return lib().el_batch_log(self._h).decode()

    def read(self, i: int):
        if not 0 <= i < len(self.scenes):
            raise ElevenError(EL_ERR_ARG, f"scene index {i} out of range (batch of {len(self.scenes)})")
        cam = self.scenes[i].camera
        rgba = np.zeros((cam.y_res, cam.x_res, 4), np.float32)
        samples = np.zeros((cam.y_res, cam.x_res), np.uint32)
        self._check(lib().el_batch_read(self._h, i, rgba.ctypes.data, samples.ctypes.data))
        return rgba, samples

    def stats(self) -> el_stats:
        s = el_stats()
        self._check(lib().el_batch_stats(self._h, C.byref(s)))
        return s

    def close(self):
        if getattr(self, "_h", None):
            lib().el_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

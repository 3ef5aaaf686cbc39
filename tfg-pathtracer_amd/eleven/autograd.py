"""torch.autograd binding of the renderer.

    image = Renderer(ctx, scene, spp, depth, seed)(params, texture)

renders the scene with per-material parameters `params` [M][6] (albedo r, g,
b, roughness, metallic, specular: the values stored in el_material) and,
optionally, the texels of one texture, and returns the mean image [H][W][3]
as a float32 CPU tensor.  Its backward runs the library's path-replay
adjoint (el_adjoint) on the accumulation the forward left in the context, so
a backward must follow its forward before the context renders again.

The renderer is deterministic for a fixed seed (Philox counters per pixel,
sample and bounce), so forward calls with the same seed use common random
numbers: finite differences of the loss are differences of the same paths.

Memory: by default nothing of a path is kept between the forward and the
backward; the adjoint recomputes every path from its RNG counters (path
replay), so the device memory of a gradient is O(pixels).  record=True keeps
each path's radiance and, in wavefront mode, every vertex (16 B per path +
16 B per vertex: O(pixels x spp x depth)), which lets the adjoint walk each
path once without tracing.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import FLAG_RECORD_PATHS, MODE_WAVEFRONT, Context, SceneData, el_material

PARAM_NAMES = ("albedo_r", "albedo_g", "albedo_b", "roughness", "metallic", "specular")


def material_params(scene: SceneData) -> np.ndarray:
    """[M][6] float32: each material's (albedo rgb, roughness, metallic, specular)."""
    out = []
    for m in scene.materials():
        out.append([m.albedo.x, m.albedo.y, m.albedo.z, m.roughness, m.metallic, m.specular])
    return np.asarray(out, np.float32)


def with_params(base: list, params: np.ndarray) -> list:
    """Copies of `base` el_material structs with the six parameters replaced."""
    out = []
    for m, p in zip(base, np.asarray(params, np.float64)):
        c = el_material()
        C.pointer(c)[0] = m
        c.albedo.x, c.albedo.y, c.albedo.z = float(p[0]), float(p[1]), float(p[2])
        c.roughness, c.metallic, c.specular = float(p[3]), float(p[4]), float(p[5])
        out.append(c)
    return out


class ParamMaterials:
    """A ctypes array of `base`'s materials whose six parameters are
    rewritten in place (through a numpy view of the array) for each call:
    the values with_params gives, without building the structs again."""

    def __init__(self, base: list):
        self.array = (el_material * len(base))(*base)
        self._view = np.ctypeslib.as_array(self.array)

    def set(self, params: np.ndarray):
        p = np.asarray(params, np.float64).astype(np.float32)   # as the c_float fields round
        v = self._view
        v["albedo"]["x"], v["albedo"]["y"], v["albedo"]["z"] = p[:, 0], p[:, 1], p[:, 2]
        v["roughness"], v["metallic"], v["specular"] = p[:, 3], p[:, 4], p[:, 5]
        return self.array


class Renderer:
    """Differentiable forward render of one uploaded scene."""

    def __init__(self, ctx: Context, scene: SceneData, spp: int, depth: int = 5, seed: int = 0,
                 grad_texture: int = -1, mode: int = MODE_WAVEFRONT, record: bool = False):
        self.ctx, self.scene = ctx, scene
        self.spp, self.depth, self.seed = spp, depth, seed
        self.grad_texture, self.mode, self.record = grad_texture, mode, record
        self.base = scene.materials()
        self._mats = ParamMaterials(self.base)

    def render(self, params: np.ndarray, texture: np.ndarray | None = None, read: bool = True) -> np.ndarray | None:
        """Forward only (numpy in, numpy out; read=False leaves the image on
        the device, e.g. for Context.residual_l2)."""
        self.ctx.set_materials(self._mats.set(params))
        if texture is not None:
            self.ctx.set_texture(self.grad_texture, np.ascontiguousarray(texture, np.float32))
        self.ctx.clear()
        # record: the paths' radiance (wavefront: also every vertex) let the
        # backward replay each path once, without tracing in wavefront mode
        self.ctx.render(self.spp, depth=self.depth, seed=self.seed, mode=self.mode,
                        flags=FLAG_RECORD_PATHS if self.record else 0)
        return self.ctx.read()[0][..., :3].copy() if read else None

    def __call__(self, params: torch.Tensor, texture: torch.Tensor | None = None) -> torch.Tensor:
        return _RenderFn.apply(params, texture, self)


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(fctx, params, texture, r: Renderer):
        tex = texture.detach().cpu().numpy() if texture is not None else None
        img = r.render(params.detach().cpu().numpy(), tex)
        fctx.r = r
        fctx.has_tex = texture is not None
        return torch.from_numpy(img)

    @staticmethod
    def backward(fctx, grad_img):
        r = fctx.r
        gt = r.grad_texture if fctx.has_tex else -1
        dmat, dtex = r.ctx.adjoint(grad_img.detach().cpu().numpy(), grad_texture=gt)
        return torch.from_numpy(dmat), (torch.from_numpy(dtex) if fctx.has_tex else None), None

"""Multi-view inverse rendering: material parameters and one albedo texture
recovered from images of several views by Adam on an L2 image loss.

One step renders every view (el_set_camera, el_render), forms dL/dI =
2 (I - target) from the mean image, runs the path-replay adjoint
(el_adjoint) and sums the gradients over the views: one complete gradient
vector (all [M][6] material parameters and the texels of the texture) per
step.  Each step and view draws its own random numbers (seed + step * views
+ view).

Decorrelation: 2 (I - T) dI/dtheta evaluated on the same paths is biased by
2 Cov(I, dI/dtheta) -- at a few samples per pixel it drives e.g. albedo
towards darker values even where E[I] = T.  By default the residual I comes
from one render and the replayed derivative from a second, independently
seeded render of the same view, which makes the gradient estimate unbiased.
With decorrelate=False the gradient is the exact derivative of the measured
loss of one render (what a finite-difference check with common random
numbers measures).  The parameters and texels are clamped to their valid
ranges after every Adam update.

Memory: by default the adjoint replays each path from its RNG counters
(O(pixels) device memory per view); record=True keeps the replayed render's
paths instead (EL_FLAG_RECORD_PATHS: O(pixels x spp x depth)).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import FLAG_RECORD_PATHS, MODE_WAVEFRONT, Context, SceneData, el_camera, make_camera
from .autograd import ParamMaterials, material_params

# seed offset of the render whose paths the adjoint replays (decorrelation)
DECORRELATE_SEED = 0x5EED0001

# valid ranges of (albedo r, g, b, roughness, metallic, specular)
PARAM_LO = np.array([0.0, 0.0, 0.0, 0.05, 0.0, 0.0], np.float32)
PARAM_HI = np.array([1.0, 1.0, 1.0, 1.0, 1.0, 1.0], np.float32)


def orbit_views(base: el_camera, n: int, x_span: float, pitches=(29.0, 35.0), res=None) -> list:
    """n cameras copied from `base`, their x positions spread evenly over
    [-x_span, x_span] and their pitch alternating over `pitches`."""
    out = []
    for k in range(n):
        xr, yr = res if res else (base.x_res, base.y_res)
        x = -x_span + 2 * x_span * k / max(1, n - 1)
        out.append(make_camera(xr, yr, position=(x, base.position.y, base.position.z),
                               rotation=(pitches[k % len(pitches)], base.rotation.y, base.rotation.z),
                               focal_length=base.focal_length, sensor_width=base.sensor_width,
                               aperture=base.aperture, focus_distance=base.focus_distance, bokeh=bool(base.bokeh)))
    return out


def survey_views(base: el_camera, extent=(51.2, 25.6), columns: int = 4, pitch: float = 60.0,
                 height: float = 22.0, res=None) -> list:
    """2 x `columns` cameras over a terrain spanning [-ex, ex] x [-ez, ez]
    (x, z): each at `height` above the ground plane, pitched down by
    `pitch` degrees and aimed at the centre of its column, the near half
    (z = -0.75 ez) or the far half (z = 0.5 ez).  With the 35 mm lens on a
    square frame (53 degrees across) and 60 degrees of pitch at height 22,
    a view spans one column of the 4 x 4 material tiles of a 102.4 x 51.2
    terrain and about half its depth, so the 8 views see every tile."""
    out = []
    ex, ez = extent
    back = height / np.tan(np.radians(pitch))
    for zc in (-0.75 * ez, 0.5 * ez):
        for i in range(columns):
            xc = -ex + (2 * i + 1) * ex / columns
            xr, yr = res if res else (base.x_res, base.y_res)
            out.append(make_camera(xr, yr, position=(xc, height, zc - back), rotation=(pitch, 0.0, 0.0),
                                   focal_length=base.focal_length, sensor_width=base.sensor_width,
                                   aperture=base.aperture, focus_distance=float(np.hypot(height, back)),
                                   bokeh=bool(base.bokeh)))
    return out


class MultiView:
    """Forward renders and summed adjoint gradients of one uploaded scene
    seen from several cameras."""

    def __init__(self, ctx: Context, scene: SceneData, cameras: list, spp: int, depth: int = 6,
                 grad_texture: int = -1, mode: int = MODE_WAVEFRONT, record: bool = False):
        self.ctx, self.scene, self.cameras = ctx, scene, list(cameras)
        self.spp, self.depth, self.grad_texture, self.mode = spp, depth, grad_texture, mode
        self.record = record
        self.base = scene.materials()
        self._mats = ParamMaterials(self.base)

    def set_params(self, params: np.ndarray, texture: np.ndarray | None = None):
        self.ctx.set_materials(self._mats.set(params))
        if texture is not None:
            self.ctx.set_texture(self.grad_texture, np.ascontiguousarray(texture, np.float32))

    def render_view(self, v: int, seed: int, spp: int | None = None, record: bool = False, read: bool = True):
        """Mean image of view v (None with read=False); with `record`, the
        render keeps its paths for a following el_adjoint (which then walks
        each path once, from the recorded vertices in wavefront mode)."""
        self.ctx.set_camera(self.cameras[v])
        self.ctx.render(spp or self.spp, depth=self.depth, seed=seed, mode=self.mode,
                        flags=FLAG_RECORD_PATHS if record else 0)
        return self.ctx.read()[0][..., :3].copy() if read else None

    def images(self, params, texture=None, seed: int = 0, spp: int | None = None) -> list:
        self.set_params(params, texture)
        return [self.render_view(v, seed + v, spp) for v in range(len(self.cameras))]

    def loss_and_grad(self, params, texture, targets, seed: int, decorrelate: bool = True):
        """Summed L2 loss over the views and its gradient (dparams [M][6],
        dtexture [H][W][3] or None)."""
        self.set_params(params, texture)
        loss = 0.0
        nv = len(self.cameras)
        dmat = dtex = None
        for v in range(nv):
            # the residual 2 (I - target) is formed on the device and stays there
            self.render_view(v, seed + v, record=self.record and not decorrelate, read=False)
            loss += self.ctx.residual_l2(targets[v])
            if decorrelate:   # the paths the adjoint replays (same view: the residual stays valid)
                self.ctx.clear()
                self.ctx.render(self.spp, depth=self.depth, seed=seed + v + DECORRELATE_SEED, mode=self.mode,
                                flags=FLAG_RECORD_PATHS if self.record else 0)
            # gradients summed over the views on the device, read once
            dmat, dtex = self.ctx.adjoint(None, grad_texture=self.grad_texture if texture is not None else -1,
                                          accumulate=v > 0, read=v == nv - 1)
        return loss, dmat, dtex


@dataclass
class InverseResult:
    params: np.ndarray
    texture: np.ndarray | None
    losses: list = field(default_factory=list)
    seconds: float = 0.0
    steps: int = 0

    @property
    def grads_per_s(self) -> float:
        return self.steps / self.seconds if self.seconds > 0 else 0.0


def block_up(c: np.ndarray, k: int) -> np.ndarray:
    """Texture of k x k texel blocks from its block values c [h][w][3] (one
    broadcast copy, on torch's threads)."""
    if k == 1:
        return c
    h, w, ch = c.shape
    src = torch.from_numpy(np.ascontiguousarray(c))
    return src[:, None, :, None, :].expand(h, k, w, k, ch).reshape(h * k, w * k, ch).numpy()


def block_sum(t: np.ndarray, k: int) -> np.ndarray:
    """Adjoint of block_up: the sum over each k x k block (rows of a block
    first, then its columns: two passes of contiguous row additions)."""
    if k == 1:
        return t
    h, w, ch = t.shape
    rows = np.ascontiguousarray(t).reshape(h // k, k, w * ch).sum(axis=1)
    return rows.reshape(h // k, w // k, k, ch).sum(axis=2)


def block_mean(t: np.ndarray, k: int) -> np.ndarray:
    return t if k == 1 else block_sum(t, k) / t.dtype.type(k * k)


def optimize(mv: MultiView, params0, texture0, targets, steps: int = 200, lr: float = 1e-2, seed: int = 1000,
             log=None, tex_blocks=((0, 1),), eps_rel: float = 0.0) -> InverseResult:
    """Adam (torch.optim) on the parameters and the texture; one gradient
    vector (all views) per step.

    eps_rel > 0: the material parameters' Adam epsilon is eps_rel times the
    RMS of the first gradient.  Adam divides a parameter's mean gradient by
    its RMS, so a parameter the images barely see (a tiny, noise-dominated
    gradient) still takes steps of about lr and random-walks; an epsilon at
    the scale of the gradients keeps such a parameter nearly still while
    parameters with gradients well above it move as before.

    tex_blocks: (first step, k) pairs, coarse to fine.  From its first step
    on, the texture is parameterized by one value per k x k texel block
    (the texture is those values repeated over their blocks; the gradient of
    a block value is the sum of its texels' gradients), starting from the
    block means of the texture so far; the optimizer state of the texture
    restarts at each change of k.  ((0, 1),) optimizes every texel."""
    p = torch.tensor(np.asarray(params0, np.float32).copy(), requires_grad=True)
    sched = sorted(tex_blocks)
    if texture0 is not None:
        h, w, _ = np.shape(texture0)
        for _, k in sched:
            if k < 1 or h % k or w % k:
                raise ValueError(f"texture block {k} does not divide the {w}x{h} texture")
    opt_p = None
    tex = None if texture0 is None else np.asarray(texture0, np.float32).copy()
    t, opt_t, kcur = None, None, None
    lo, hi = torch.from_numpy(PARAM_LO), torch.from_numpy(PARAM_HI)
    res = InverseResult(params=None, texture=None)
    t0 = time.perf_counter()
    for k in range(steps):
        if tex is not None:
            kk = [b for s0, b in sched if s0 <= k][-1]
            if kk != kcur:   # a new (finer) parameterization from the texture so far
                cur = tex if t is None else block_up(t.detach().numpy(), kcur)
                t = torch.tensor(block_mean(cur, kk).astype(np.float32), requires_grad=True)
                opt_t, kcur = torch.optim.Adam([t], lr=lr), kk
            tex = block_up(t.detach().numpy(), kcur)
        loss, dmat, dtex = mv.loss_and_grad(p.detach().numpy(), tex, targets, seed + k * len(mv.cameras))
        if opt_p is None:
            eps = eps_rel * float(np.sqrt(np.mean(np.square(dmat, dtype=np.float64)))) if eps_rel > 0 else 1e-8
            opt_p = torch.optim.Adam([p], lr=lr, eps=max(eps, 1e-8))
        opt_p.zero_grad()
        p.grad = torch.from_numpy(dmat.astype(np.float32))
        opt_p.step()
        if t is not None:
            opt_t.zero_grad()
            t.grad = torch.from_numpy(block_sum(np.asarray(dtex, np.float32), kcur))
            opt_t.step()
        with torch.no_grad():
            p.copy_(torch.max(torch.min(p, hi), lo))
            if t is not None:
                t.clamp_(0.0, 1.0)
        res.losses.append(loss)
        if log and (k % 20 == 0 or k == steps - 1):
            log(f"inverse step {k}: loss {loss:.6g}")
    res.seconds = time.perf_counter() - t0
    res.steps = steps
    res.params = p.detach().numpy().copy()
    res.texture = None if t is None else block_up(t.detach().numpy(), kcur).copy()
    return res


def param_rmse(a: np.ndarray, b: np.ndarray, mask: np.ndarray | None = None, weights=None) -> float:
    """RMSE of a - b over `mask`; with `weights` (e.g. |dL/dtheta| at the
    start: the error along the directions the images see) a weighted RMSE."""
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2
    w = np.ones_like(d) if weights is None else np.abs(np.asarray(weights, np.float64))
    if mask is not None:
        d, w = d[mask], w[mask]
    return float(np.sqrt((w * d).sum() / w.sum()))


__all__ = ["MultiView", "InverseResult", "optimize", "orbit_views", "survey_views", "param_rmse", "material_params",
           "block_up", "block_sum", "block_mean", "PARAM_LO", "PARAM_HI"]

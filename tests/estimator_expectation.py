"""Exact expectations of the reference estimator on single-plane scenes.

Test infrastructure: computes E[L] of renderingKernel's estimator
(src/tfg-pathtracer/kernel.cu:175-258 pointLight / hdriLight, 339-358 shade,
404-480 path loop, clamp and NaN rejection) for a camera looking at one
planar quad, by exact enumeration and quadrature instead of Monte Carlo, so
that the oracle's composed renderer (or_render) can be checked against it.
The composition below is written from kernel.cu and HDRI.hpp directly; it
shares no code with the oracle's path walk.  It uses the oracle only for
primitives that tests/test_oracle_pins.py pins on their own: DisneyEval /
DisneyPdf / DisneySample (closed forms, reciprocity, quadrature, chi^2),
the HDRI CDF (closed form) and the camera ray (closed form).

Random numbers: the oracle draws u = (w >> 8) / 2^24 from Philox words, so
each uniform takes the 2^24 values i / 2^24 with equal probability.
  * The HDRI pixel k = binarySearch(cdf, r1, W*H) (HDRI.hpp:130-142, which
    is not a lower bound) and the point light floor(count * r1), with the
    SAME r1 (kernel.cu:347-348): their joint distribution is enumerated over
    all 2^24 values of r1.
  * The BRDF direction (DisneySample r1 r2 r3) is integrated by the midpoint
    rule over (r1, r2) per lobe, the lobe chosen with probability
    0.5 (1 - metallic) (Disney.hpp:162).
  * The camera jitter is integrated by the midpoint rule over the pixel.

One vertex (kernel.cu:339-358), with h = HDRI pdf of the sampled pixel,
p = count / (2 pi) and c = DisneyPdf of the sampled BRDF direction:
    L = T (w1 Hl + w2 Pl + w3 B),  w = (h, p, c) / (h + p + c)
    Hl = f(wk) |N.wk| env_k / h [shadow ray clear]
    Pl = rad / d^2 f(l) |N.l| / p [shadow ray clear]
    B  = Le f(b) |N.b| / c          (Le: the emission of THIS vertex)
and the next vertex starts from T f(b) |N.b| / c; a ray that escapes adds
env(-dir) T (kernel.cu:414-419).  The sample is clamped to [0, 10] per
channel and dropped if any channel is NaN (kernel.cu:447-449).

Every function takes a `variant` naming one plausible misreading of the
reference, so a test can show that its tolerance separates the reference
semantics from each of them.
"""
from __future__ import annotations

import math

import numpy as np

import eleven
import oracle_py as orc

N24 = 1 << 24


# ---------------------------------------------------------------- scene ---
def plane_scene(normal, albedo=(0.7, 0.5, 0.3), roughness=0.6, metallic=0.0, specular=0.5, emission=(0, 0, 0),
                env=None, lights=(), res=(4, 4), half=60.0, cam_pos=(0.0, 5.0, -10.0), sensor=0.002):
    """A square of side 2*half through the origin with unit normal `normal`
    (two triangles, vertex normals = the face normal), one material, the
    environment image `env` [H][W][3], point lights [(position, radiance)],
    and a narrow pinhole camera at cam_pos looking at the origin."""
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 0.0, 1.0])
    u = np.cross(n, a)
    u /= np.linalg.norm(u)
    v = np.cross(n, u)
    c = [half * (sx * u + sy * v) for sx, sy in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    tris = [c[0], c[1], c[2], c[0], c[2], c[3]]
    V = np.array(tris, np.float32).reshape(2, 9)
    N = np.tile(n, 6).astype(np.float32).reshape(2, 9)
    mat = eleven.make_material(albedo=albedo, roughness=roughness, metallic=metallic, specular=specular,
                               emission=emission)
    p = np.asarray(cam_pos, np.float64)
    dx = -p
    pitch = math.degrees(math.atan2(-dx[1], math.hypot(dx[0], dx[2])))
    yaw = math.degrees(math.atan2(dx[0], dx[2]))
    cam = dict(x_res=res[0], y_res=res[1], position=tuple(cam_pos), rotation=(pitch, yaw, 0.0),
               sensor_width=sensor)
    env = np.full((32, 64, 3), 0.5, np.float32) if env is None else np.asarray(env, np.float32)
    s = eleven.SceneData.from_arrays(V, cam, normals=N, object_id=np.zeros(2, np.int32), mesh_material=[0],
                                     materials=[mat], hdri=env, lights=list(lights))
    s.plane_normal = n.astype(np.float32).astype(np.float64)
    s.env_img = env
    return s


# ------------------------------------------------------------------ HDRI ---
def binary_search_all(cdf: np.ndarray, length: int) -> np.ndarray:
    """HDRI.hpp:130-142 evaluated for every r = i / 2^24 (float32): the pixel
    index each r selects."""
    r = (np.arange(N24, dtype=np.float64) / N24).astype(np.float32)
    frm = np.zeros(N24, np.int64)
    to = np.full(N24, length - 1, np.int64)
    res = np.full(N24, -1, np.int64)
    act = to - frm > 0
    while act.any():
        idx = np.nonzero(act)[0]
        f, t, rv = frm[idx], to[idx], r[idx]
        m = f + (t - f) // 2
        cm = cdf[m]
        eq = rv == cm
        res[idx[eq]] = m[eq]
        lt, gt = rv < cm, rv > cm
        to[idx] = np.where(lt, m - 1, t)
        frm[idx] = np.where(gt, m + 1, f)
        act[idx] = (~eq) & (to[idx] - frm[idx] > 0)
    left = res < 0
    res[left] = to[left]
    return res


def binary_search(cdf: np.ndarray, r: np.ndarray, length: int) -> np.ndarray:
    """HDRI.hpp:130-142 for an array of float32 values r."""
    r = np.asarray(r, np.float32)
    frm = np.zeros(r.shape, np.int64)
    to = np.full(r.shape, length - 1, np.int64)
    res = np.full(r.shape, -1, np.int64)
    act = to - frm > 0
    while act.any():
        m = frm + (to - frm) // 2
        cm = cdf[np.clip(m, 0, len(cdf) - 1)]
        eq = act & (r == cm)
        res[eq] = m[eq]
        to = np.where(act & (r < cm), m - 1, to)
        frm = np.where(act & (r > cm), m + 1, frm)
        act = act & ~eq & (to - frm > 0)
    return np.where(res < 0, to, res)


def pixel_light_joint(env: np.ndarray, n_lights: int, variant=None):
    """P[k, li]: probability that one shade() draws HDRI pixel k and point
    light li from its single r1 (kernel.cu:185, 236; HDRI.hpp:154-162), r1
    uniform over the 2^24 values i / 2^24.  Every comparison the search makes
    is against a CDF value, so k is constant for r strictly between two
    consecutive distinct CDF values; the light index floor(count * r) is
    constant between multiples of 1 / count.  Each open interval between
    consecutive break points is evaluated once at a representative value and
    weighted by the grid values inside it; grid values equal to a break point
    are evaluated on their own (exact, and equal to enumerating all 2^24
    values: see binary_search_all).  variant 'independent_light_pick' draws
    the light from its own number."""
    H, W = env.shape[:2]
    cdf, rsum = orc.hdri_cdf(env)
    nl = max(1, n_lights)
    L = W * H
    pts = np.unique(np.concatenate([cdf[:L].astype(np.float64), np.arange(1, nl) / nl, [0.0, 1.0]]))
    pts = pts[(pts >= 0) & (pts <= 1)]
    g = pts * N24                                   # grid coordinates of the break points (exact)
    # grid values strictly inside (pts[j], pts[j+1]): i in (g_j, g_{j+1})
    lo = np.floor(g[:-1]).astype(np.int64) + 1
    hi = np.ceil(g[1:]).astype(np.int64) - 1
    cnt = np.maximum(0, np.minimum(hi, N24 - 1) - lo + 1)
    P = np.zeros((L, nl))
    inner = cnt > 0
    rep = ((lo[inner] + np.minimum(hi[inner], N24 - 1)) // 2).astype(np.float64) / N24
    k = binary_search(cdf, rep, L)
    li = np.minimum((nl * rep).astype(np.int64), nl - 1)
    np.add.at(P, (k, li), cnt[inner])
    # grid values that are break points
    on = (g == np.floor(g)) & (g < N24)
    r0 = g[on] / N24
    np.add.at(P, (binary_search(cdf, r0, L), np.minimum((nl * r0).astype(np.int64), nl - 1)), 1)
    P /= N24
    if variant == "independent_light_pick":
        P = P.sum(1, keepdims=True) * np.full((1, nl), 1.0 / nl)
    return P, cdf, rsum


def env_pixel_dirs(W: int, H: int) -> np.ndarray:
    """Direction of HDRI pixel k's corner: -reverseSphericalMapping(x/W, y/H)
    normalised (kernel.cu:238-244, Texture.hpp:195-207), [W*H][3]."""
    y, x = np.divmod(np.arange(W * H), W)
    iu = (x / W).astype(np.float64)
    iv = (y / H).astype(np.float64)
    phi, theta = iu * 2 * np.pi, iv * np.pi
    px, py, pz = np.cos(phi - np.pi), -np.cos(theta), -np.sin(phi - np.pi)
    a = np.sqrt(np.maximum(0.0, 1 - py * py))
    d = np.stack([a * px, py, a * pz], -1)
    return -d / np.linalg.norm(d, axis=-1, keepdims=True)


def env_pdf(env: np.ndarray, rsum: float, variant=None) -> np.ndarray:
    """HDRI::pdf of every pixel (HDRI.hpp:145-152): lum / sum * W * H /
    (2 pi sin(y pi / H)); infinite or NaN on row 0.  variant 'equirect_pdf'
    uses the equirectangular Jacobian 2 pi^2 sin instead."""
    H, W = env.shape[:2]
    lum = env.reshape(-1, 3).astype(np.float64).sum(-1)
    y = np.arange(W * H) // W
    den = 2.0 * np.pi * np.sin(y / H * np.pi)
    if variant == "equirect_pdf":
        den = den * np.pi
    with np.errstate(divide="ignore", invalid="ignore"):
        return (lum / rsum) * W * H / den


def env_lookup(env: np.ndarray, d: np.ndarray) -> np.ndarray:
    """Escaped ray: nearest texel at sphericalMapping(-d) (kernel.cu:416-417,
    Texture.hpp:144-156, limitUV), d [..., 3] -> [..., 3]."""
    H, W = env.shape[:2]
    p = -np.asarray(d, np.float64)
    p = p / np.linalg.norm(p, axis=-1, keepdims=True)
    theta = np.arccos(np.clip(-p[..., 1], -1, 1))
    phi = np.arctan2(-p[..., 2], p[..., 0]) + np.pi
    u, v = phi / (2 * np.pi), theta / np.pi
    u = u - (u > 1) - (u < 0)
    v = v - (v > 1) - (v < 0)
    x = np.clip((u * W).astype(np.int64), 0, W - 1)
    y = np.clip((v * H).astype(np.int64), 0, H - 1)
    return env[y % H, x % W].astype(np.float64)


# ------------------------------------------------------------------ BRDF ---
def hitdata_of(scene, normal):
    m = scene.materials()[0]
    return orc.hitdata(normal=tuple(normal), albedo=(m.albedo.x, m.albedo.y, m.albedo.z),
                       emission=(m.emission.x, m.emission.y, m.emission.z),
                       metallic=float(m.metallic) ** 2.2, roughness=float(m.roughness) ** 2.2,
                       specular=m.specular)


def brdf_samples(hd, d, n=96):
    """Midpoint quadrature of DisneySample over (r1, r2) per lobe: weights q,
    directions b (not normalised, as sampled), f(b), c = DisneyPdf(b),
    |N.b|."""
    metallic = hd[0]
    ratio = 0.5 * (1.0 - metallic)
    N = hd[18:21]
    g = (np.arange(n) + 0.5) / n
    out = []
    for lobe_p, r3 in ((ratio, 0.0), (1.0 - ratio, 1.0 - 1e-9)):
        if lobe_p <= 0:
            continue
        for r1 in g:
            for r2 in g:
                b = orc.disney_sample(hd, d, r1, r2, r3)
                out.append((lobe_p / (n * n), b, orc.disney_eval(hd, d, b), orc.disney_pdf(hd, d, b),
                            abs(float(np.dot(b, N)))))
    q = np.array([o[0] for o in out])
    B = np.array([o[1] for o in out])
    F = np.array([o[2] for o in out])
    C = np.array([o[3] for o in out])
    cos = np.array([o[4] for o in out])
    return q, B, F, C, cos


def pixel_rays(scene, x, y, m=3):
    """Camera rays of kernel pixel (x, y) at an m x m midpoint grid of the
    jitter (r1, r2) of calculateCameraRay (kernel.cu:263-277)."""
    g = (np.arange(m) + 0.5) / m
    out = []
    for r1 in g:
        for r2 in g:
            o, d = orc.camera_ray(scene.view.camera, x, y, np.array([r1, r2, 0.5, 0.5, 0.5], np.float32))
            out.append((o.astype(np.float64), d.astype(np.float64)))
    return out


# ------------------------------------------------------------ expectation --
VARIANTS = ("swap_w1_w2", "equirect_pdf", "emission_unweighted", "no_escape", "escape_sign",
            "independent_light_pick", "no_clamp", "no_nan_reject")


class PixelTerms:
    """Everything the estimator's expectation at kernel pixel (x, y) of a
    plane_scene depends on, for each camera ray of the jitter grid: per HDRI
    pixel f(wk) |N.wk| env_k, per light rad / d^2 f(l) |N.l|, and the BRDF
    quadrature (weights, f(b), DisneyPdf(b), |N.b|, escape value)."""

    _joint = {}

    def __init__(self, scene, x, y, jitter=3, nq=96):
        self.scene = scene
        env = scene.env_img
        H, W = env.shape[:2]
        N = scene.plane_normal
        self.lights = [(np.array([l.position.x, l.position.y, l.position.z], np.float64),
                        np.array([l.radiance.x, l.radiance.y, l.radiance.z], np.float64))
                       for l in (scene.view.lights[i] for i in range(scene.view.light_count))]
        dirs = env_pixel_dirs(W, H)
        val = env.reshape(-1, 3).astype(np.float64)
        hd = hitdata_of(scene, N)
        self.Le = hd[12:15]
        self.rays = []
        for o, d in pixel_rays(scene, x, y, jitter):
            t = float(np.dot(-o, N) / np.dot(d, N))
            pos = o + t * d
            fk = np.array([orc.disney_eval(hd, d, w) for w in dirs])
            Hn = fk * np.abs(dirs @ N)[:, None] * val
            Pn = np.zeros((max(1, len(self.lights)), 3))
            for i, (lp, lr) in enumerate(self.lights):
                ld = lp - pos
                dist = np.linalg.norm(ld)
                ld = ld / dist
                Pn[i] = lr / dist ** 2 * orc.disney_eval(hd, d, ld) * abs(float(np.dot(ld, N)))
            q, B, F, C, cosb = brdf_samples(hd, d, nq)
            self.rays.append(dict(Hn=Hn, Pn=Pn, q=q, B=B, F=F, C=C, cosb=cosb))

    def joint(self, variant):
        env = self.scene.env_img
        key = (env.tobytes(), len(self.lights), variant == "independent_light_pick")
        if key not in PixelTerms._joint:
            PixelTerms._joint[key] = pixel_light_joint(env, len(self.lights),
                                                      "independent_light_pick" if key[2] else None)
        return PixelTerms._joint[key]

    def expectation(self, depth=1, variant=None):
        """(E[clamped L | accepted], P(accepted)) under the reference
        semantics or one misreading (VARIANTS):
          swap_w1_w2             the HDRI and point-light MIS weights exchanged
          equirect_pdf           HDRI pdf with 2 pi^2 sin(theta)
          emission_unweighted    emission added as Le T at the hit instead of w3 B
          no_escape              escaped rays add nothing
          escape_sign            escaped rays read env(dir) instead of env(-dir)
          independent_light_pick the light chosen by its own random number
          no_clamp / no_nan_reject
        depth 1: one vertex; depth 2: the vertex and the escape of its
        extension ray (the plane is the only surface).

        With h_k infinite (row 0) or zero (a black pixel) w1 Hl is NaN (inf /
        inf, or 0 * 0 / 0), whatever the visibility and the other numbers, so
        such a draw rejects the sample.  Otherwise L is
        (Hn_k + Pn_l + Bn_b) / (h_k + p + c_b) + T_b env_b, which the sums
        below take term by term (sum over b of q_b / (h_k + p + c_b) per
        distinct h_k) when no sample can reach the clamp, and sample by
        sample otherwise."""
        env = self.scene.env_img
        H, W = env.shape[:2]
        nl = len(self.lights)
        P, _, rsum = self.joint(variant)
        h = env_pdf(env, rsum, "equirect_pdf" if variant == "equirect_pdf" else None)
        good = np.isfinite(h) & (h > 0)
        p_l = nl / (2.0 * np.pi) if nl else 0.0
        Le = self.Le
        Pg = P[good]                                    # [K, nl]
        hg = h[good]
        p_bad = float(P[~good].sum())
        acc_sum = np.zeros(3)
        for r in self.rays:
            Hn, Pn, q, B, F, C, cosb = r["Hn"], r["Pn"], r["q"], r["B"], r["F"], r["C"], r["cosb"]
            Hg = Hn[good]
            Bn = Le[None, :] * F * cosb[:, None]                       # c * B
            esc = np.zeros(3)
            escb = np.zeros((len(q), 3))
            if depth >= 2 and variant != "no_escape":
                bd = B / np.linalg.norm(B, axis=-1, keepdims=True)
                e = env_lookup(env, -bd if variant == "escape_sign" else bd)
                escb = (F * cosb[:, None] / C[:, None]) * e            # T * env(-dir)
                esc = (q[:, None] * escb).sum(0)
            cH = (p_l / hg) if variant == "swap_w1_w2" else np.ones_like(hg)
            cP = (hg / p_l) if (variant == "swap_w1_w2" and nl) else np.ones_like(hg)
            bound = ((cH[:, None] * Hg).max() + (cP.max() * Pn.max() if nl else 0) + Bn.max()) / (
                hg.min() + p_l + C.min()) + escb.max()
            if bound < 10.0 or variant == "no_clamp":
                hu, inv = np.unique(hg, return_inverse=True)
                G = (q[None, :] / (hu[:, None] + p_l + C[None, :])).sum(1)[inv]      # [K]
                s = ((Pg.sum(1) * G * cH)[:, None] * Hg).sum(0)
                for li in range(nl):
                    s += (Pg[:, li] * G * cP).sum() * Pn[li]
                if variant == "emission_unweighted":
                    s += Pg.sum() * Le
                else:
                    Kb = (Pg.sum(1)[:, None] / (hu[inv][:, None] + p_l + C[None, :])).sum(0)   # [B]
                    s += (q * Kb) @ Bn
                s += Pg.sum() * esc
                acc_sum += s
            else:
                for k0 in range(0, len(hg), 128):
                    ks = slice(k0, min(len(hg), k0 + 128))
                    D = hg[ks][:, None] + p_l + C[None, :]
                    for li in range(max(1, nl)):
                        num = (cH[ks][:, None] * Hg[ks])[:, None, :] + Bn[None, :, :]
                        if nl:
                            num = num + (cP[ks][:, None] * Pn[li][None, :])[:, None, :]
                        L = num / D[..., None]
                        if variant == "emission_unweighted":
                            L = L - Bn[None, :, :] / D[..., None] + Le[None, None, :]
                        L = np.clip(L + escb[None, :, :], 0.0, 10.0)
                        acc_sum += ((Pg[ks, li][:, None] * q[None, :])[..., None] * L).sum((0, 1))
        acc_sum /= len(self.rays)
        p_acc = 1.0 if variant == "no_nan_reject" else 1.0 - p_bad
        return acc_sum / p_acc, p_acc


def oracle_estimate(scene, x, y, spp, depth, seed=99, threads=0):
    """or_render (FP64 shading and accumulation, the reference's shadow rule)
    of kernel pixel (x, y): mean of the accepted samples, its standard
    error, the accepted count and the sample count."""
    rb = orc.RefBVH(scene.vertices())
    s, ss, cnt, _, _ = orc.render(scene, rb, spp, depth=depth, seed=seed, precision=1, shadow_mode=0,
                                  threads=threads, window=(x, x + 1, y, y + 1))
    H = scene.view.camera.y_res
    row = H - 1 - y
    n = float(cnt[row, x])
    mean = s[row, x] / max(n, 1.0)
    var = np.maximum(ss[row, x] / max(n, 1.0) - mean ** 2, 0.0) * n / max(n - 1, 1.0)
    return mean, np.sqrt(var / max(n, 1.0)), n, spp

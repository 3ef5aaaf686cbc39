"""Pins of the CPU oracle (CPU only): known-answer vectors, closed forms,
brute force and distribution tests, each chosen so that a plausible mistake
(a dropped term, a swapped operand, a wrong sign or index) fails.
"""
import math
import struct

import numpy as np
import pytest

import oracle_py as orc


# ------------------------------------------------------------- Philox -----
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    """Random123 known-answer vectors for philox4x32_10."""
    assert orc.philox(ctr, key).tolist() == expect


# ------------------------------------------------------ Moller-Trumbore -----
TRI = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)


@pytest.mark.parametrize("o,d,expect", [
    ((0.25, 0.25, 1), (0, 0, -1), (True, 1.0, 0.25, 0.25)),   # interior hit
    ((0.8, 0.8, 1), (0, 0, -1), (False,)),                     # u + v > 1
    ((0.5, 0.0, 1), (0, 0, -1), (True, 1.0, 0.5, 0.0)),        # on an edge (inclusive)
    ((1.0, 0.0, 1), (0, 0, -1), (True, 1.0, 1.0, 0.0)),        # on a vertex (inclusive)
    ((0.25, 0.25, 1), (1, 0, 0), (False,)),                    # parallel: det = 0
    ((0.25, 0.25, -1), (0, 0, -1), (False,)),                  # behind the origin
    ((0.25, 0.25, 0), (0, 0, -1), (True, 0.0, 0.25, 0.25)),    # t = 0 accepted
    ((-0.1, 0.2, 1), (0, 0, -1), (False,)),                    # u < 0
])
def test_tri_hit_special_cases(o, d, expect):
    h, t, u, v = orc.tri_hit(TRI, o, d)
    assert h == expect[0]
    if h:
        assert (t, u, v) == expect[1:]


def test_tri_hit_small_determinant_rejected():
    # |det| = |e1 . (d x e2)| = 1e-8 < 1e-7: rejected even though the ray passes through
    tri = np.array([0, 0, 0, 1e-4, 0, 0, 0, 1e-4, 0], np.float32)
    h, *_ = orc.tri_hit(tri, (2e-5, 2e-5, 1), (0, 0, -1))
    assert not h
    h, *_ = orc.tri_hit(tri * 1e2, (2e-3, 2e-3, 1), (0, 0, -1))   # det = 1e-4: accepted
    assert h


def test_tri_hit_barycentric_invariant():
    """For every hit, o + t d and v0 + u e1 + v e2 name the same point, and t
    agrees with an FP64 plane intersection."""
    rng = np.random.default_rng(0)
    n = 4000
    tris = rng.uniform(-1, 1, size=(n, 9)).astype(np.float32)
    hits = 0
    for i in range(n):
        v = tris[i].astype(np.float64).reshape(3, 3)
        target = v[0] + rng.uniform(0, 1) * (v[1] - v[0]) * 0.5 + rng.uniform(0, 1) * (v[2] - v[0]) * 0.5
        o = rng.uniform(-3, 3, 3)
        d = (target - o) / np.linalg.norm(target - o)
        o32, d32 = o.astype(np.float32), d.astype(np.float32)
        h, t, u, w = orc.tri_hit(tris[i], o32, d32)
        if not h:
            continue
        hits += 1
        p_ray = o32.astype(np.float64) + t * d32.astype(np.float64)
        p_bar = v[0] + u * (v[1] - v[0]) + w * (v[2] - v[0])
        nrm = np.cross(v[1] - v[0], v[2] - v[0])
        # FP32 error grows with the conditioning |e1||e2| / |det| of the hit
        kappa = np.linalg.norm(v[1] - v[0]) * np.linalg.norm(v[2] - v[0]) / abs(np.dot(d32, nrm))
        tol = 2e-6 * (1 + np.abs(o).max() + t) * kappa
        assert np.linalg.norm(p_ray - p_bar) < tol
        t64 = np.dot(v[0] - o32, nrm) / np.dot(d32.astype(np.float64), nrm)
        assert abs(t - t64) < tol
        assert u >= 0 and w >= 0 and u + w <= 1
    assert hits > 0.9 * n


def test_closest_brute_rules():
    near = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    far = [0, 0, -1, 1, 0, -1, 0, 1, -1]
    v = np.array([far, near, near], np.float32)          # index 1 and 2 coincide
    rays = np.array([[0.2, 0.2, 1, 0, 0, -1], [5, 5, 1, 0, 0, -1], [0.2, 0.2, -2, 0, 0, 1]], np.float32)
    tri, t, u, w = orc.closest_brute(v, rays)
    assert tri.tolist() == [1, -1, 0]                      # nearer wins, ties -> lower index
    assert t[0] == 1.0 and np.isinf(t[1]) and t[2] == 1.0


def test_reference_bvh_equals_brute_force():
    """The depth-18 / 14-bin tree restatement finds the brute-force closest hit
    for random (non-degenerate) rays on random soups."""
    rng = np.random.default_rng(1)
    for n in (1, 37, 2000):
        c = rng.uniform(-1, 1, size=(n, 1, 3))
        v = (c + rng.normal(scale=0.2, size=(n, 3, 3))).astype(np.float32).reshape(n, 9)
        ref = orc.RefBVH(v, depth=10 if n < 100 else 18)
        o = rng.uniform(-2, 2, size=(5000, 3))
        d = rng.normal(size=(5000, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        rays = np.concatenate([o, d], 1).astype(np.float32)
        a, at = ref.closest(rays)
        b, bt, _, _ = orc.closest_brute(v, rays)
        assert np.array_equal(a, b)
        assert np.array_equal(at[a >= 0], bt[b >= 0])
        tm = rng.uniform(0, 4, 5000).astype(np.float32)
        assert np.array_equal(ref.any(rays, tm), (b >= 0) & (bt < tm))
    assert orc.lib().or_refbvh_node_count(orc.RefBVH(v, depth=18).h) == 2 ** 19 - 1


# ------------------------------------------------------------- Disney -----
def _hd(**kw):
    return orc.hitdata(**kw)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("albedo,metallic,rough,spec", [
    ((0.8, 0.4, 0.1), 0.0, 0.3, 0.5), ((0.2, 0.6, 0.9), 1.0, 0.05, 0.5), ((0.5, 0.5, 0.5), 0.4, 1.0, 0.2)])
def test_disney_eval_normal_incidence_closed_form(prec, albedo, metallic, rough, spec):
    """V = L = N: f = albedo (1-m)/pi + Cspec0 / (4 pi a^2), a = max(0.001, roughness)."""
    hd = _hd(normal=(0, 0, 1), albedo=albedo, metallic=metallic, roughness=rough, specular=spec)
    f = orc.disney_eval(hd, (0, 0, -1), (0, 0, 1), prec)
    a = max(0.001, rough)
    cspec0 = (1 - metallic) * spec * 0.08 + metallic * np.array(albedo)
    expect = np.array(albedo) * (1 - metallic) / math.pi + cspec0 / (4 * math.pi * a * a)
    np.testing.assert_allclose(f, expect, rtol=1e-5 if prec == 0 else 1e-12)


def test_disney_eval_reciprocity_and_hemispheres():
    rng = np.random.default_rng(2)
    n = np.array([0.3, 0.8, 0.52])
    n /= np.linalg.norm(n)
    for _ in range(300):
        hd = _hd(normal=n, albedo=rng.uniform(0, 1, 3), metallic=rng.uniform(), roughness=rng.uniform(0.05, 1),
                 specular=rng.uniform(), sheen=rng.uniform(), subsurface=rng.uniform(), clearcoat=rng.uniform(),
                 clearcoatGloss=rng.uniform())
        a, b = rng.normal(size=3), rng.normal(size=3)
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        if a @ n < 0:
            a = -a
        if b @ n < 0:
            b = -b
        f_ab = orc.disney_eval(hd, -a, b, 1)
        f_ba = orc.disney_eval(hd, -b, a, 1)
        np.testing.assert_allclose(f_ab, f_ba, rtol=1e-9, atol=1e-12)
        assert np.all(f_ab >= 0)
        assert np.all(orc.disney_eval(hd, -a, -b, 1) == 0)        # L below the surface
        assert orc.disney_pdf(hd, -a, -b, 1) == 1.0               # reference returns 1 there


def _mirror(theta):
    """V and L mirrored about N = +z at angle theta in the xz-plane (H = N,
    L.H = cos theta); returns the ray direction (-V) and L."""
    c, s_ = math.cos(theta), math.sin(theta)
    return np.array([s_, 0.0, -c]), np.array([s_, 0.0, c])


@pytest.mark.parametrize("metallic", [0.0, 0.3])
@pytest.mark.parametrize("tint", [0.0, 1.0])
def test_disney_sheen_lobe_closed_form(metallic, tint):
    """Sheen adds FH sheen Csheen (1 - m) (Disney.hpp:236): at theta = 60 deg
    mirrored, L.H = 1/2 so FH = (1/2)^5 = 1/32; Csheen = 1 without tint and
    albedo / (0.3 r + 0.6 g + 0.1 b) with full tint."""
    alb = np.array([0.6, 0.3, 0.2])
    d, L = _mirror(math.pi / 3)
    base = orc.disney_eval(_hd(albedo=alb, metallic=metallic, sheen=0.0, sheenTint=tint), d, L, 1)
    with_ = orc.disney_eval(_hd(albedo=alb, metallic=metallic, sheen=0.7, sheenTint=tint), d, L, 1)
    csheen = alb / (0.3 * 0.6 + 0.6 * 0.3 + 0.1 * 0.2) if tint else np.ones(3)
    np.testing.assert_allclose(with_ - base, 0.7 * csheen * (1 - metallic) / 32, rtol=1e-12)
    # normal incidence: FH = 0, no sheen at all
    n0 = orc.disney_eval(_hd(albedo=alb, metallic=metallic, sheen=0.7, sheenTint=tint), (0, 0, -1), (0, 0, 1), 1)
    n1 = orc.disney_eval(_hd(albedo=alb, metallic=metallic, sheen=0.0, sheenTint=tint), (0, 0, -1), (0, 0, 1), 1)
    np.testing.assert_allclose(n0, n1, rtol=1e-14)


def test_disney_subsurface_lobe_closed_form():
    """Subsurface replaces Fd by ss in the diffuse term (Disney.hpp:213-222,
    238), so f(ss=1) - f(ss=0) = (ss - Fd) Cdlin (1 - m) / pi.  Hand-worked
    at roughness 1/2: normal incidence (FL = FV = 0): Fd = 1, Fss = 1,
    ss = 1.25 (1 (1/2 - 1/2) + 1/2) = 0.625; theta = 60 deg mirrored
    (N.L = N.V = L.H = 1/2, FL = FV = 1/32): Fd90 = 0.75, Fd = (1 - 0.25/32)^2
    = 0.98443603515625, Fss90 = 0.125, Fss = (1 - 0.875/32)^2 =
    0.9460601806640625, ss = 1.25 (Fss (1 - 1/2) + 1/2) =
    1.2162876129150390625."""
    alb, m = np.array([0.8, 0.5, 0.3]), 0.25
    for (d, L), expect in [(((0, 0, -1), (0, 0, 1)), 0.625 - 1.0),
                           (_mirror(math.pi / 3), 1.2162876129150390625 - 0.98443603515625)]:
        f1 = orc.disney_eval(_hd(albedo=alb, metallic=m, roughness=0.5, subsurface=1.0), d, L, 1)
        f0 = orc.disney_eval(_hd(albedo=alb, metallic=m, roughness=0.5, subsurface=0.0), d, L, 1)
        np.testing.assert_allclose((f1 - f0) * math.pi / (alb * (1 - m)), expect, rtol=1e-12)


@pytest.mark.parametrize("gloss", [0.0, 0.5, 1.0])
def test_disney_clearcoat_lobe_closed_form(gloss):
    """Clearcoat adds 0.25 cc Gr Fr Dr (Disney.hpp:232-238): Dr = GTR1 with
    a = 0.1 + (0.001 - 0.1) gloss, which at N.H = 1 is (a^2 - 1) / (pi
    ln(a^2) a^2); Gr = G(N.L) G(N.V) with G(x) = 1 / (x + sqrt(1/16 + x^2 -
    x^2 / 16)); Fr = 0.04 + 0.96 FH.  Normal incidence: Gr = 1/4, Fr = 0.04;
    theta = 60 deg mirrored: FH = 1/32, Fr = 0.07."""
    a = 0.1 + (0.001 - 0.1) * gloss
    dr = (a * a - 1) / (math.pi * math.log(a * a) * a * a)
    g = lambda x: 1.0 / (x + math.sqrt(0.0625 + x * x - 0.0625 * x * x))
    for (d, L), gr, fr in [(((0, 0, -1), (0, 0, 1)), 0.25, 0.04), (_mirror(math.pi / 3), g(0.5) ** 2, 0.07)]:
        f1 = orc.disney_eval(_hd(clearcoat=0.6, clearcoatGloss=gloss), d, L, 1)
        f0 = orc.disney_eval(_hd(clearcoat=0.0, clearcoatGloss=gloss), d, L, 1)
        # 1e-9: at gloss 1 (a = 0.001) the oracle's lerp rounds a differently in FP64
        np.testing.assert_allclose(f1 - f0, 0.25 * 0.6 * gr * fr * dr, rtol=1e-9)


def test_disney_specular_tint_closed_form():
    """Cspec0 = lerp(spec 0.08 lerp(1, Ctint, st), Cdlin, m) (Disney.hpp:207):
    at normal incidence Fs = Cspec0, Gs = 1/4 and Ds = 1 / (pi a^2), so
    f(st=1) - f(st=0) = spec 0.08 (Ctint - 1) (1 - m) / (4 pi a^2)."""
    alb, m, spec, r = np.array([0.9, 0.4, 0.1]), 0.2, 0.6, 0.4
    ctint = alb / (0.3 * 0.9 + 0.6 * 0.4 + 0.1 * 0.1)
    f1 = orc.disney_eval(_hd(albedo=alb, metallic=m, specular=spec, roughness=r, specularTint=1.0),
                         (0, 0, -1), (0, 0, 1), 1)
    f0 = orc.disney_eval(_hd(albedo=alb, metallic=m, specular=spec, roughness=r, specularTint=0.0),
                         (0, 0, -1), (0, 0, 1), 1)
    np.testing.assert_allclose(f1 - f0, spec * 0.08 * (ctint - 1) * (1 - m) / (4 * math.pi * r * r), rtol=1e-12)


@pytest.mark.parametrize("aniso", [0.3, 0.9])
def test_disney_anisotropic_lobe_against_isotropic(aniso):
    """GTR2_aniso with H in the tangent plane's x-z section equals GTR2 with
    a = ax scaled by ax / ay (1 + (a^2-1) h_z^2 = a^2 (h_x^2 / a^2 + h_z^2)),
    and the anisotropic Smith term there equals the isotropic one at a = ax;
    so with metallic 1 (no diffuse or sheen) and V = N, f(aniso, r) for L in
    the T-N plane equals f(0, ax) ax / ay, and in the B-N plane f(0, ay)
    ay / ax, with aspect = sqrt(1 - 0.9 aniso), ax = r / aspect, ay = r
    aspect (Disney.hpp:224-231).  The basis for N = +z is T = +x, B = +y
    (createBasis: up x N).  At normal incidence anisotropy changes nothing."""
    r = 0.3
    aspect = math.sqrt(1 - 0.9 * aniso)
    ax, ay = r / aspect, r * aspect
    alb = np.array([0.9, 0.7, 0.4])
    hd = lambda an, rough: _hd(albedo=alb, metallic=1.0, roughness=rough, anisotropic=an)
    for th in (0.3, 0.8, 1.2):
        Lt = np.array([math.sin(th), 0.0, math.cos(th)])
        Lb = np.array([0.0, math.sin(th), math.cos(th)])
        np.testing.assert_allclose(orc.disney_eval(hd(aniso, r), (0, 0, -1), Lt, 1),
                                   orc.disney_eval(hd(0.0, ax), (0, 0, -1), Lt, 1) * ax / ay, rtol=1e-11)
        np.testing.assert_allclose(orc.disney_eval(hd(aniso, r), (0, 0, -1), Lb, 1),
                                   orc.disney_eval(hd(0.0, ay), (0, 0, -1), Lb, 1) * ay / ax, rtol=1e-11)
    np.testing.assert_allclose(orc.disney_eval(hd(aniso, r), (0, 0, -1), (0, 0, 1), 1),
                               orc.disney_eval(hd(0.0, r), (0, 0, -1), (0, 0, 1), 1), rtol=1e-12)


def _hemisphere_pdf_mass(hd, dirv, nb_theta=12, nb_phi=8, sub=24):
    """Integral of DisneyPdf over (cos theta, phi) bins of the upper hemisphere (midpoint rule)."""
    mass = np.zeros((nb_theta, nb_phi))
    for i in range(nb_theta):
        for j in range(nb_phi):
            acc = 0.0
            for a in range(sub):
                ct = (i + (a + 0.5) / sub) / nb_theta
                st = math.sqrt(1 - ct * ct)
                for b in range(sub // 2):
                    ph = 2 * math.pi * (j + (b + 0.5) / (sub // 2)) / nb_phi
                    L = np.array([st * math.cos(ph), st * math.sin(ph), ct])
                    acc += orc.disney_pdf(hd, dirv, L, 1)
            mass[i, j] = acc / (sub * sub // 2) * (1.0 / nb_theta) * (2 * math.pi / nb_phi)
    return mass


@pytest.mark.parametrize("metallic,rough", [(0.0, 0.6), (0.5, 0.3), (1.0, 0.5)])
def test_disney_sample_matches_pdf(metallic, rough):
    """Histogram of DisneySample directions vs the integral of DisneyPdf per bin
    (non-degenerate basis, clearcoat = anisotropic = 0, where the pdf is the
    sampler's density)."""
    n = np.array([0.0, 0.0, 1.0])
    hd = _hd(normal=n, metallic=metallic, roughness=rough)
    v = np.array([0.4, 0.1, 0.91])
    v /= np.linalg.norm(v)
    dirv = -v
    rng = np.random.default_rng(3)
    N = 60000
    nb_t, nb_p = 12, 8
    counts = np.zeros((nb_t, nb_p))
    below = 0
    for _ in range(N):
        L = orc.disney_sample(hd, dirv, rng.uniform(), rng.uniform(), rng.uniform(), 1)
        L /= np.linalg.norm(L)
        if L[2] <= 0:
            below += 1
            continue
        ph = math.atan2(L[1], L[0]) % (2 * math.pi)
        counts[min(nb_t - 1, int(L[2] * nb_t)), min(nb_p - 1, int(ph / (2 * math.pi) * nb_p))] += 1
    mass = _hemisphere_pdf_mass(hd, dirv, nb_t, nb_p)
    expect = mass * N
    ok = expect > 20
    chi2 = ((counts[ok] - expect[ok]) ** 2 / expect[ok]).sum()
    dof = ok.sum()
    assert chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)
    assert abs((N - below) / N - mass.sum()) < 0.01


def test_disney_sample_degenerate_basis_goes_straight_up():
    """N = +Y: createBasis gives T = B = 0, so a diffuse sample is N * sqrt(1 - u1)."""
    hd = _hd(normal=(0, 1, 0), metallic=0.0, roughness=0.5)
    for u1, u2 in [(0.1, 0.3), (0.64, 0.9), (0.0, 0.5)]:
        L = orc.disney_sample(hd, (0.3, -0.9, 0.1), u1, u2, 0.0, 1)   # r3 = 0 < diffuse ratio
        np.testing.assert_allclose(L, [0, math.sqrt(1 - u1), 0], atol=1e-12)


def test_ggx_and_cosine_sampling_closed_forms():
    rng = np.random.default_rng(4)
    for _ in range(500):
        a, r1, r2 = rng.uniform(0.01, 1), rng.uniform(), rng.uniform()
        h = orc.importance_sample_ggx(a, r1, r2)
        theta = math.atan(a * math.sqrt(r2 / (1 - r2)))       # textbook GGX inversion
        np.testing.assert_allclose(h, [math.sin(theta) * math.cos(2 * math.pi * r1),
                                       math.sin(theta) * math.sin(2 * math.pi * r1), math.cos(theta)], atol=1e-9)
        c = orc.cosine_sample_hemisphere(r1, r2)               # Malley: cos(theta) = sqrt(1 - u1)
        np.testing.assert_allclose(c, [math.sqrt(r1) * math.cos(2 * math.pi * r2),
                                       math.sqrt(r1) * math.sin(2 * math.pi * r2), math.sqrt(1 - r1)], atol=1e-12)
    # lens sampler: folded u2+u3 is triangular -> uniform over the disk (P(r <= x) = x^2)
    r = np.array([math.hypot(*orc.uniform_circle(*rng.uniform(size=3))) for _ in range(20000)])
    xs = np.linspace(0.05, 0.95, 19)
    emp = np.array([(r <= x).mean() for x in xs])
    assert np.max(np.abs(emp - xs ** 2)) < 0.015


# --------------------------------------------------------------- HDRI -----
def test_hdri_cdf_closed_form():
    rgb = np.array([[[1, 0, 0], [0, 2, 0]], [[0, 0, 0], [3, 0, 1]]], np.float32)
    cdf, s = orc.hdri_cdf(rgb)
    assert s == 7.0
    np.testing.assert_allclose(cdf, [0, 1 / 7, 3 / 7, 3 / 7, 1.0], rtol=1e-6)


@pytest.mark.parametrize("r,expect", [(0.1, 0), (0.3, 1), (0.6, 3), (0.8, 3), (0.5, 2), (0.25, 1), (0.0, 0)])
def test_hdri_binary_search_hand_cases(r, expect):
    """HDRI.hpp:130-142 traced by hand on cdf = [0, .25, .5, .75, 1], length 4:
    not a lower bound (r = 0.6 lies in [0.5, 0.75) but returns 3)."""
    cdf = np.array([0, 0.25, 0.5, 0.75, 1.0], np.float32)
    assert orc.hdri_search(cdf, r, 4) == expect


def test_hdri_search_monotone_and_off_by_at_most_one():
    rng = np.random.default_rng(5)
    for _ in range(40):
        L = int(rng.integers(2, 300))
        lum = rng.uniform(0, 1, L) * (rng.uniform(size=L) < 0.8)
        lum[0] += 1e-3
        cdf = np.concatenate([[0], np.cumsum(lum / lum.sum())]).astype(np.float32)
        rs = np.sort(rng.uniform(size=400)).astype(np.float32)
        res = np.array([orc.hdri_search(cdf, r, L) for r in rs])
        assert np.all(np.diff(res) >= 0)
        ideal = np.searchsorted(cdf[:L], rs, side="right") - 1
        assert np.all((res - ideal >= -1) & (res - ideal <= 1))


def test_hdri_pdf_closed_form():
    rng = np.random.default_rng(6)
    rgb = rng.uniform(0, 2, size=(8, 16, 3)).astype(np.float32)
    cdf, s = orc.hdri_cdf(rgb)
    for x, y in [(3, 1), (15, 7), (0, 4)]:
        lum = rgb[y, x].astype(np.float64).sum()
        expect = lum / s * 16 * 8 / (2 * math.pi * math.sin(y / 8 * math.pi))
        assert abs(orc.hdri_pdf(rgb, 0.0, s, x, y) - expect) <= 1e-5 * expect


# ------------------------------------------------------------- camera -----
def test_camera_pinhole_closed_form():
    import eleven
    cam = eleven.make_camera(200, 100, position=(1, 2, 3))
    for x, y in [(0, 0), (100, 50), (199, 99)]:
        o, d = orc.camera_ray(cam, x, y, [0.5, 0.5, 0, 0, 0])
        sw, sh, f = 0.035, 0.035 * 0.5, 0.035
        v = np.array([x / 200 * sw - sw / 2, y / 100 * sh - sh / 2, f])
        np.testing.assert_allclose(o, [1, 2, 3])
        # the reference adds the camera position to the sensor coordinate and
        # subtracts it again (kernel.cu:263-303): FP32 cancellation ~ ulp(3)/f
        np.testing.assert_allclose(d, v / np.linalg.norm(v), atol=2e-5)
    # +90 degrees about x maps the forward axis (0,0,1) to (0,-1,0)
    cam = eleven.make_camera(10, 10, rotation=(90, 0, 0))
    o, d = orc.camera_ray(cam, 5, 5, [0.5, 0.5, 0, 0, 0])
    np.testing.assert_allclose(d, [0, -1, 0], atol=1e-6)


def test_camera_defocus_focus_point():
    """With defocus every lens sample of a pixel passes through the same focus point."""
    import eleven
    cam = eleven.make_camera(64, 64, focus_distance=10.0, aperture=0.05, bokeh=True)
    o0, d0 = orc.camera_ray(eleven.make_camera(64, 64), 20, 30, [0.3, 0.7, 0, 0, 0])
    focus = o0.astype(np.float64) + d0 * (10.0 + 0.035)
    rng = np.random.default_rng(7)
    for _ in range(20):
        o, d = orc.camera_ray(cam, 20, 30, [0.3, 0.7, *rng.uniform(size=3)])
        to_f = focus - o
        np.testing.assert_allclose(np.cross(d, to_f / np.linalg.norm(to_f)), 0, atol=5e-6)
        assert np.linalg.norm(o[:2]) <= 0.035 / 0.05 * 0.5 + 1e-6 and o[2] == 0


def test_disney_eval_grad_matches_central_differences():
    """The analytic derivative the adjoint uses (albedo rgb, roughness,
    metallic, specular) against central differences of DisneyEval in FP64,
    over random materials (all lobes on), random tilted normals and random
    V, L in the upper hemisphere."""
    rng = np.random.default_rng(11)
    keys = [("albedo", 0), ("albedo", 1), ("albedo", 2), ("roughness", None), ("metallic", None),
            ("specular", None)]
    worst = 0.0
    for _ in range(400):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        kw = dict(metallic=rng.uniform(0.05, 0.95), roughness=rng.uniform(0.05, 1.0), specular=rng.uniform(0, 1),
                  specularTint=rng.uniform(), sheenTint=rng.uniform(), sheen=rng.uniform(),
                  subsurface=rng.uniform(), clearcoat=rng.uniform(), clearcoatGloss=rng.uniform(),
                  anisotropic=rng.uniform(0, 0.9))
        albedo = rng.uniform(0.05, 0.95, 3)
        a, b = rng.normal(size=3), rng.normal(size=3)
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        if a @ n < 0.05:
            a = a - (a @ n - 0.3) * n
            a /= np.linalg.norm(a)
        if b @ n < 0.05:
            b = b - (b @ n - 0.3) * n
            b /= np.linalg.norm(b)
        hd = _hd(normal=n, albedo=albedo, **kw)
        g = orc.disney_eval_grad(hd, -a, b)
        for k, (name, comp) in enumerate(keys):
            h = 1e-6
            if name == "albedo":
                al_p, al_m = albedo.copy(), albedo.copy()
                al_p[comp] += h
                al_m[comp] -= h
                fp = orc.disney_eval(_hd(normal=n, albedo=al_p, **kw), -a, b, 1)
                fm = orc.disney_eval(_hd(normal=n, albedo=al_m, **kw), -a, b, 1)
            else:
                kp, km = dict(kw), dict(kw)
                kp[name] += h
                km[name] -= h
                fp = orc.disney_eval(_hd(normal=n, albedo=albedo, **kp), -a, b, 1)
                fm = orc.disney_eval(_hd(normal=n, albedo=albedo, **km), -a, b, 1)
            fd = (fp - fm) / (2 * h)
            err = np.abs(g[k] - fd) / (np.abs(fd) + 1e-3 * np.abs(orc.disney_eval(hd, -a, b, 1)).max() + 1e-9)
            worst = max(worst, float(err.max()))
    assert worst < 1e-5, worst


def test_disney_eval_grad_zero_outside_hemisphere():
    hd = _hd(normal=(0.2, 0.9, 0.3), albedo=(0.3, 0.5, 0.7), metallic=0.4, roughness=0.3)
    n = np.array([0.2, 0.9, 0.3]) / np.linalg.norm([0.2, 0.9, 0.3])
    assert np.all(orc.disney_eval_grad(hd, -n, -n) == 0)

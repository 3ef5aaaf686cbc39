"""Pins the oracle's composed estimator (or_render) against exact expectations.

tests/estimator_expectation.py computes E[L] of the reference estimator
(kernel.cu:175-258, 339-358, 404-480) for one pixel of a single-plane scene
by enumeration and quadrature, from the reference text and the separately
pinned primitives, sharing no code with the oracle's path walk.  Each case
renders the pixel with or_render (FP64, the reference's shadow rule) at
65,536-262,144 samples and requires:
  * every channel of the mean of the accepted samples within 4 standard
    errors of the expectation, and the accepted-sample count within 4
    binomial standard deviations of P(accepted);
  * every misreading listed for the case (a swapped MIS weight, the HDRI pdf
    with the equirectangular 2 pi^2 Jacobian -- a factor pi --, unweighted
    emission, a dropped or mirrored escape term, an independent light pick,
    no clamp, no NaN rejection) at least 8 standard errors away, so the check
    would fail if the oracle made that mistake.

Cases:
  a  tilted diffuse plane, constant environment, depth 1 (w1 Hl; HDRI pdf
     1 / (2 pi sin theta), HDRI.hpp:151; the pole row's infinite pdf rejects
     the sample)
  b  the same plane with one point light and a dim environment (w2 Pl,
     pdf count / 2 pi, kernel.cu:182)
  b2 two lights and a gradient sky: the light pick floor(count r1) shares r1
     with the HDRI pick (kernel.cu:185, 347-348)
  c  an emissive plane: w3 B with the emission of the vertex itself (Q9,
     kernel.cu:349)
  d  a +Y plane at depth 2: the degenerate tangent frame of createBasis
     (Disney.hpp:41-45, Q1) and the escape term env(-dir) T
     (kernel.cu:414-419)
  e  a pole-row sun, black pixels, a bright sun and a strong light: the
     [0, 10] clamp and the NaN rejection (kernel.cu:447-449)
"""
import numpy as np
import pytest

import estimator_expectation as ee
import oracle_py as orc

TILT = (0.3, 0.8, -0.5)


def grad_env(W=64, H=32, top=1.2, bottom=0.2):
    v = (np.arange(H) + 0.5) / H
    col = top + (bottom - top) * v
    return np.stack([col, 0.9 * col, 0.7 * col], -1)[:, None, :].repeat(W, 1).astype(np.float32)


def clamp_env():
    env = np.full((32, 64, 3), 0.3, np.float32)
    env[0] = (6.0, 5.5, 5.0)                 # pole row: infinite pdf
    env[12:14, 20:24] = 0.0                  # black pixels: zero pdf
    env[7:9, 40:42] = (400.0, 380.0, 300.0)  # a bright sun
    return env


CASES = {
    "a_constant_env": (lambda: ee.plane_scene(TILT, env=np.full((32, 64, 3), 0.8, np.float32)), (1, 2), 1, 65536,
                       ["equirect_pdf", "swap_w1_w2", "no_nan_reject"]),
    "b_point_light": (lambda: ee.plane_scene(TILT, env=np.full((32, 64, 3), 0.05, np.float32),
                                             lights=[((2.0, 4.0, -1.0), (20.0, 20.0, 20.0))]), (2, 1), 1, 65536,
                      ["swap_w1_w2", "equirect_pdf"]),
    "b2_shared_light_pick": (lambda: ee.plane_scene(TILT, env=grad_env(),
                                                    lights=[((2.0, 4.0, -1.0), (30.0, 30.0, 30.0)),
                                                            ((-3.0, 2.0, 1.0), (5.0, 8.0, 12.0))]), (1, 1), 1, 65536,
                             ["independent_light_pick", "swap_w1_w2"]),
    "c_emission_at_vertex": (lambda: ee.plane_scene(TILT, env=np.full((32, 64, 3), 0.3, np.float32),
                                                    emission=(2.0, 1.0, 0.5)), (2, 2), 1, 65536,
                             ["emission_unweighted"]),
    "d_up_plane_escape": (lambda: ee.plane_scene((0.0, 1.0, 0.0), env=grad_env()), (1, 2), 2, 65536,
                          ["no_escape", "escape_sign", "equirect_pdf"]),
    "e_clamp_and_nan": (lambda: ee.plane_scene(TILT, env=clamp_env(), roughness=0.4,
                                               lights=[((1.0, 2.5, -1.0), (250.0, 200.0, 150.0))]), (1, 2), 1,
                        262144, ["no_clamp", "no_nan_reject", "swap_w1_w2"]),
}


def test_joint_pixel_light_distribution_equals_full_enumeration():
    """The interval form of the (HDRI pixel, light) distribution equals the
    search run on every one of the 2^24 values r1 takes, and the vectorised
    search equals the oracle's (pinned) binarySearch."""
    rng = np.random.default_rng(3)
    env = rng.uniform(0, 1, size=(4, 8, 3)).astype(np.float32)
    env[0, 1] = 0.0
    env[1, 2] = 0.0                       # zero-width CDF steps
    nl = 3
    P, cdf, _ = ee.pixel_light_joint(env, nl)
    k = ee.binary_search_all(cdf, 32)
    r = np.arange(ee.N24) / ee.N24
    li = np.minimum((nl * r).astype(np.int64), nl - 1)
    full = np.bincount(k * nl + li, minlength=32 * nl).reshape(32, nl) / ee.N24
    assert np.array_equal(P, full)
    rs = rng.uniform(0, 1, 3000).astype(np.float32)
    rs[:32] = cdf[:32]                    # values equal to CDF entries
    got = ee.binary_search(cdf, rs, 32)
    assert all(got[i] == orc.hdri_search(cdf, float(rs[i]), 32) for i in range(len(rs)))


@pytest.mark.parametrize("case", list(CASES))
def test_or_render_matches_exact_expectation(case):
    make, (x, y), depth, spp, wrong = CASES[case]
    scene = make()
    terms = ee.PixelTerms(scene, x, y, jitter=3, nq=64)
    E, p_acc = terms.expectation(depth)
    mean, se, n, N = ee.oracle_estimate(scene, x, y, spp, depth)
    z = (mean - E) / se
    zc = (n / N - p_acc) / np.sqrt(p_acc * (1 - p_acc) / N)
    print(f"{case}: E {E} oracle {mean} z {z} accepted {n:.0f}/{N} (expected {p_acc:.5f}, z {zc:.2f})")
    assert np.all(np.abs(z) < 4), f"oracle mean off the expectation: z {z}"
    assert abs(zc) < 4, f"accepted-sample count off: z {zc}"
    assert p_acc < 1.0                     # every case draws the pole row
    for v in wrong:
        Ev, pv = terms.expectation(depth, v)
        zv = np.abs((mean - Ev) / se).max()
        zcv = abs(n / N - pv) / np.sqrt(max(pv * (1 - pv), 1e-12) / N)
        print(f"  misreading {v}: max |z| {zv:.1f}, count z {zcv:.1f}")
        assert max(zv, zcv) > 8, f"the check cannot tell the reference from misreading {v}"

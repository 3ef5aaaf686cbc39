"""Statistics of an image against a high-sample reference image.

The GPU image g is the mean of n accepted samples per pixel of the same
estimator as the reference R (the mean of n_ref samples, with per-sample
variance v).  Under the hypothesis that both estimate the same expectation,
each channel difference g - R has variance v / n + v / n_ref.

relMSE (relative mean squared error) = mean over pixels and channels of
(g - R)^2 / (R^2 + 1e-2).  Its expectation under the hypothesis is the same
mean of (v / n + v / n_ref) / (R^2 + 1e-2), and its standard deviation, with
each difference Gaussian (n >= 256 samples) and the three channels of a pixel
counted as fully correlated (they share one path), is
sqrt(sum_p 2 (sum_c e_pc)^2) / N with e_pc the expected term.  The checks:
  * the measured relMSE within 3 standard deviations of the predicted one
    (consistent with the sample variance at 3 sigma);
  * the summed difference over the image, sum_p sum_c (g - R), within 3 of
    its standard deviation (a global bias test, sensitive to ~1e-3
    relative);
  * at most max(5, 5e-4 N) pixel channels with |z| > 5 (outliers a mean
    would hide).

Rare large samples (a light or sun hit clamped to 10) make single pixels
heavy-tailed: the reference's per-sample variance v, estimated from n_ref
samples, can miss an event that a GPU render of n samples happens to draw,
and the GPU mean then sits many of the reference's standard deviations
away although it is right (checked: the oracle with the GPU's random
numbers reproduces such pixels exactly).  So the GPU image is rendered as
independent batches (batched_render) and, for the per-pixel z only, the
variance of its mean is the larger of v / n and the spread of the batch
means / batches.  The relMSE and bias tests keep v / n + v / n_ref, whose
expectation is exact (the larger of two estimates is biased upwards).
"""
import numpy as np


def batched_render(ctx, spp, batches=16, **kw):
    """spp samples as `batches` successive progressive renders (the same
    samples as one render of spp); returns the mean image [H][W][3], the
    accepted counts and the variance of the mean from the batch spread."""
    prev_s, prev_n, means = 0.0, 0.0, []
    for _ in range(batches):
        ctx.render(spp // batches, **kw)
        img, cnt, _, _ = ctx.read()
        s = img[..., :3].astype(np.float64) * cnt[..., None]
        c = cnt.astype(np.float64)
        means.append((s - prev_s) / np.maximum(c - prev_n, 1.0)[..., None])
        prev_s, prev_n = s, c
    m = np.stack(means)
    return img[..., :3], cnt, m.var(axis=0, ddof=1) / batches


def compare(g, n, R, v, n_ref, g_var=None):
    """g, R, v [..., 3]; n (accepted samples of g) and n_ref (of R) per pixel
    [...] or scalars; g_var (optional, [..., 3]): the variance of g from its
    own batches."""
    g = np.asarray(g, np.float64)
    R = np.asarray(R, np.float64)
    v = np.asarray(v, np.float64)
    n = np.asarray(n, np.float64)[..., None]
    nr = np.asarray(n_ref, np.float64)
    if nr.ndim:
        nr = nr[..., None]
    var = v / np.maximum(n, 1.0) + v / np.maximum(nr, 1.0)
    var_z = var if g_var is None else np.maximum(v / np.maximum(n, 1.0), np.asarray(g_var, np.float64)) + \
        v / np.maximum(nr, 1.0)
    w = 1.0 / (R ** 2 + 1e-2)
    d = g - R
    N = d.size
    measured = float(np.mean(d ** 2 * w))
    e = var * w
    predicted = float(e.mean())
    sd = float(np.sqrt(2.0 * (e.sum(-1) ** 2).sum()) / N)
    sdp = np.sqrt(var).sum(-1)                                # per pixel, channels fully correlated
    zg = float(d.sum() / np.sqrt((sdp ** 2).sum())) if sdp.any() else 0.0
    ok = var > 0          # pixels the reference saw vary; the others must match exactly
    z = np.zeros_like(d)
    z[ok] = d[ok] / np.sqrt(var_z[ok])
    det_bad = int(np.sum(np.abs(d[~ok]) > 1e-5 * (np.abs(R[~ok]) + 1e-3)))
    return {"relmse": measured, "relmse_predicted": predicted, "relmse_sd": sd,
            "relmse_z": (measured - predicted) / sd if sd > 0 else 0.0, "bias_z": zg,
            "max_abs_z": float(np.abs(z).max()), "n_z5": int(np.sum(np.abs(z) > 5)), "n": N,
            "deterministic_mismatch": det_bad,
            "worst": [(tuple(int(i) for i in np.unravel_index(k, d.shape)), float(g.flat[k]), float(R.flat[k]),
                       float(z.flat[k])) for k in np.argsort(-np.abs(z).ravel())[:3]]}


def assert_consistent(stats, label=""):
    s = stats
    print(f"{label} relMSE {s['relmse']:.4e} predicted {s['relmse_predicted']:.4e} (z {s['relmse_z']:.2f}); "
          f"bias z {s['bias_z']:.2f}; max |z| {s['max_abs_z']:.2f}, |z| > 5: {s['n_z5']} of {s['n']}; "
          f"largest (index, value, reference, z): {s['worst']}", flush=True)
    assert s["deterministic_mismatch"] == 0, f"{label}: zero-variance pixels differ"
    assert abs(s["relmse_z"]) < 3, f"{label}: relMSE not consistent with the sample variance at 3 sigma"
    assert abs(s["bias_z"]) < 3, f"{label}: summed difference beyond 3 sigma"
    assert s["n_z5"] <= max(5, 5e-4 * s["n"]), f"{label}: too many pixels beyond 5 sigma"

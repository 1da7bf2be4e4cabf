"""Per-pixel statistics of (w, X, Y, Z) films with sums of squared XYZ, for
comparing two renders of the same job (box filter: every sample has weight 1
in exactly one pixel, up to measure-zero boundary samples)."""
import numpy as np


# Film sums are float32: E[x^2] - E[x]^2 cannot resolve a variance below a
# few dozen ulps of E[x^2] (pixels whose samples are nearly equal, e.g. sky
# seen directly), so the variance is floored there.
VAR_FLOOR_ULPS = 64.0 * np.finfo(np.float32).eps


def mean_se(img, sq):
    w = img[..., 0:1]
    n = np.maximum(w, 1.0)
    mean = img[..., 1:] / n
    ex2 = sq / n
    var = np.maximum(ex2 - mean * mean, VAR_FLOOR_ULPS * ex2)
    se = np.sqrt(var / np.maximum(n - 1.0, 1.0))
    return mean, se


def compare(img_a, sq_a, img_b, sq_b, floor=1e-6):
    """Returns (fraction of pixel channels outside 3 sigma, mean z, max |z|)."""
    ma, sa = mean_se(img_a, sq_a)
    mb, sb = mean_se(img_b, sq_b)
    sig = np.sqrt(sa * sa + sb * sb) + floor * (np.abs(ma) + np.abs(mb) + floor)
    z = (ma - mb) / sig
    return float((np.abs(z) > 3.0).mean()), float(z.mean()), float(np.abs(z).max())


def rmse(img, ref):
    m = img[..., 1:] / np.maximum(img[..., 0:1], 1.0)
    r = ref[..., 1:] / np.maximum(ref[..., 0:1], 1.0)
    return float(np.sqrt(((m - r) ** 2).mean()))

"""Source pictures of the reference's hierarchical fixtures, restated so the golden frame-0 coefficients can be
recomputed: picture() is jpeg/examples/make_hier.rs:12-32 and shrink() is jpeg/src/hierarchical.rs:283-303.
Test helper only."""
import math

import numpy as np


def _round_half_away(x):
    f = math.floor(x)
    return f + 1 if x - f >= 0.5 else f


def picture(w, h, channels):
    """8-bit picture(w, h, Gray|Rgb, 8) as (h, w) or (h, w, 3) uint8."""
    out = np.zeros((h, w, channels), np.uint8)
    for y in range(h):
        for x in range(w):
            for c in range(channels):
                smooth = (float(x) * (3.0 + float(c)) + float(y) * 2.0) / float(w + h)
                edge = 0.25 if (x // 7 + y // 5) % 3 == 0 else 0.0
                tex = float((x * 31 + y * 17 + c * 5) % 13) / 130.0
                s = smooth + edge + tex
                fr = s - math.trunc(s)
                out[y, x, c] = _round_half_away(fr * 255.0)
    return out[:, :, 0] if channels == 1 else out


def shrink(img):
    """Halve width and height with a rounded box average over the pixels that exist (per channel)."""
    a = img.astype(np.int64)
    if a.ndim == 2:
        a = a[:, :, None]
    h, w, ch = a.shape
    nw, nh = -(-w // 2), -(-h // 2)
    out = np.zeros((nh, nw, ch), np.int64)
    for y in range(nh):
        for x in range(nw):
            ys = [yy for yy in (2 * y, 2 * y + 1) if yy < h]
            xs = [xx for xx in (2 * x, 2 * x + 1) if xx < w]
            n = len(ys) * len(xs)
            s = sum(a[yy, xx] for yy in ys for xx in xs)
            out[y, x] = (s + n // 2) // n
    out = out.astype(np.uint8)
    return out[:, :, 0] if img.ndim == 2 else out

"""Host restatement of the predicted threshold of gmax_mode 2 (spec_band in
csrc/qp3d_kernels.hpp), for reporting how many steps of a run kept their prediction: the
update of step k takes ts = rd(0.01 Gp), Gp extrapolated from G_{k-1} and G_{k-2} (ratio
clamped to [1/2, 2]), and marks the cells whose |grad u| lies in (ts (1 - 1/32), ts (1 + 1/32)];
the step kept its prediction when rd(0.01 G_k) lies in that band."""
import numpy as np

SPEC_BAND = 1.0 / 32


def s_threshold(G):
    """rd(0.01 * G) in float32: the largest float not above 0.01*G (double product)."""
    d = 0.01 * float(G)
    f = np.float32(d)
    if float(f) > d:
        f = np.nextafter(f, np.float32(-np.inf))
    return np.float32(f)


def spec_band(gm, k, band=SPEC_BAND):
    """(predicted G, ts, lo, hi) of step k from the per-step maxima gm (float32)."""
    gp = np.float32(0)
    if k >= 1:
        g1 = np.float32(gm[k - 1])
        gp = g1
        if k >= 2:
            g2 = np.float32(gm[k - 2])
            if g2 > 0 and g1 > 0 and g1 <= 3.4e38:
                r = np.float32(min(max(np.float32(g1 / g2), np.float32(0.5)), np.float32(2.0)))
                gp = np.float32(g1 * r)
    ts = s_threshold(gp)
    return gp, ts, np.float32(ts * np.float32(1 - band)), np.float32(ts * np.float32(1 + band))


def spec_valid_steps(gm):
    """Per step: True when rd(0.01 G_k) lies in the band marked by the prediction."""
    out = []
    for k in range(len(gm)):
        _, _, lo, hi = spec_band(gm, k)
        out.append(bool(lo <= s_threshold(gm[k]) <= hi))
    return np.array(out, bool)

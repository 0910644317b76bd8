"""Host restatement of the predicted threshold of gmax_mode 2 (spec_band and predict_next in
csrc/qp3d_kernels.hpp). The update of step k thresholds at ts = rd(0.01 Gp) with Gp the
prediction of G_k, and marks the cells whose |grad u| lies in (ts (1 - 1/128), ts (1 + 1/128)];
the step kept its prediction when rd(0.01 G_k) lies in that band. The prediction is an AR(6)
fit of log G over the last 32 steps (the ratio rule before that); predict_G restates it in
float64 for the oracle tests (not bit for bit: the device's own predictions are read with
Plan.step_predictions)."""
import numpy as np

SPEC_BAND = 1.0 / 128
AR_P, AR_W = 6, 32


def s_threshold(G):
    """rd(0.01 * G) in float32: the largest float not above 0.01*G (double product)."""
    d = 0.01 * float(G)
    f = np.float32(d)
    if float(f) > d:
        f = np.nextafter(f, np.float32(-np.inf))
    return np.float32(f)


def predict_G(gm, k):
    """Predicted G_k from gm[0..k-1] (0 for k = 0)."""
    if k < 1:
        return np.float32(0)
    gk = float(gm[k - 1])
    gp = gk
    done = False
    n0 = k - AR_W - AR_P
    win = np.asarray(gm[max(n0, 0):k], np.float64)
    if n0 >= 0 and gk > 0 and np.all(win > 0) and np.all(win <= 3.4e38):
        lg = np.log(win) - np.log(gk)
        rows = np.array([[lg[AR_P + j - 1 - m] for m in range(AR_P)] + [1.0] for j in range(AR_W)])
        y = lg[AR_P:AR_P + AR_W]
        M = rows.T @ rows
        M += 1e-9 * np.trace(M) / 7 * np.eye(7)
        try:
            c = np.linalg.solve(M, rows.T @ y)
            L = c[AR_P] + sum(c[m] * lg[AR_P + AR_W - 1 - m] for m in range(AR_P))
            if np.isfinite(L) and abs(L) < 50:
                gp, done = gk * np.exp(L), True
        except np.linalg.LinAlgError:
            pass
    if not done and k >= 2:
        g1 = float(gm[k - 2])
        if g1 > 0 and gk > 0:
            gp = gk * min(max(gk / g1, 0.5), 2.0)
    if gk > 0:
        gp = min(max(gp, 0.5 * gk), 2.0 * gk)
    return np.float32(gp)


def band(gp, width=SPEC_BAND):
    """(ts, lo, hi) of a predicted G, in float32 as the kernels form them."""
    ts = s_threshold(gp)
    return ts, np.float32(ts * np.float32(1 - width)), np.float32(ts * np.float32(1 + width))


def steps_in_band(gm, gpred):
    """Per step: True when rd(0.01 G_k) lies in the band of the step's predicted G."""
    out = []
    for k in range(len(gm)):
        _, lo, hi = band(gpred[k] if k >= 1 else np.float32(0))
        out.append(bool(lo <= s_threshold(gm[k]) <= hi))
    return np.array(out, bool)

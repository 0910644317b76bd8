"""The host restatement of the gmax_mode 2 prediction (qp3d/threshold.py), which the GPU
test compares the device's stored predictions against, pinned by what the mathematics fixes:
a log-G series made of a linear trend and one sinusoid obeys an exact linear recurrence of
order 4, so an AR(6) fit with a constant term predicts it to rounding; and rd(0.01 G) is the
largest float not above 0.01 G."""
import numpy as np

from conftest import band, predict_G, s_threshold, steps_in_band


def test_ar_prediction_is_exact_on_a_trend_plus_sinusoid():
    k = np.arange(200)
    g = np.exp(-7.7 - 0.002 * k + 0.02 * np.sin(2 * np.pi * k / 5.2 + 0.3)).astype(np.float32)
    err = [abs(float(predict_G(g, j)) / float(g[j]) - 1) for j in range(38, 200)]
    assert max(err) < 1e-6
    # before 38 steps of history: the ratio rule G_{k-1} * (G_{k-1} / G_{k-2})
    assert predict_G(g, 0) == 0 and predict_G(g, 1) == g[0]
    r = float(g[9]) / float(g[8])
    assert abs(float(predict_G(g, 10)) / (float(g[9]) * r) - 1) < 1e-6


def test_threshold_rounds_down_and_band_brackets_it():
    rng = np.random.default_rng(3)
    for G in rng.uniform(1e-6, 1e3, 200).astype(np.float32):
        t = s_threshold(G)
        assert float(t) <= 0.01 * float(G) < float(np.nextafter(t, np.float32(np.inf)))
        ts, lo, hi = band(G)
        assert ts == t and lo < ts < hi and abs(float(hi) / float(ts) - 1 - 1 / 128) < 1e-6
    gm = np.array([0, 1.0, 1.0, 1.1], np.float32)
    # step 0 is in band only if G_0 = 0; step 1 predicted exactly; step 3 off by 10%
    assert list(steps_in_band(gm, np.array([0, 1.0, 1.0, 1.0], np.float32))) == [True, True, True, False]

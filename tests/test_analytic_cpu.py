"""The trajectory post-processing helper against closed forms (CPU)."""
import numpy as np

import analytic as A


def _damped(T, amp=0.9, lam=0.3, period=2.0):
    w = 2 * np.pi / period
    return amp * np.exp(-lam * T) * np.cos(w * T), lam, w


def test_oscillation_extrema_closed_form():
    """yn = a e^(-lam T) cos(w T): the extrema lie at
    T_k = (k pi - atan(lam / w)) / w, alternate in sign and shrink."""
    ny, R, U0 = 34, 4.0, 0.01
    H = ny - 2
    T = np.linspace(0.0, 10.0, 20001)
    yn, lam, w = _damped(T)
    Y = 0.5 * (ny - 1) + yn * (0.5 * H - R)
    t = T * H / U0
    Tg, yng, idx = A.oscillation_extrema(t, Y, ny, R, U0, prominence=0.05)
    assert np.allclose(Tg, T) and np.allclose(yng, yn)
    k = np.arange(1, 40)
    Tk = (k * np.pi - np.arctan(lam / w)) / w
    Tk = Tk[(Tk > 0) & (Tk < 9.5)]  # one at the series' end has no prominence
    assert len(idx) == len(Tk)
    assert np.allclose(T[idx], Tk, atol=2 * (T[1] - T[0]))
    assert np.all(np.diff(np.sign(yn[idx])) != 0)
    assert np.all(np.diff(np.abs(yn[idx])) < 0)
    # t_skip drops the extrema at or before it
    _, _, late = A.oscillation_extrema(t, Y, ny, R, U0, t_skip=1.5, prominence=0.05)
    assert np.array_equal(late, idx[T[idx] > 1.5])


def test_oscillation_extrema_prominence_filters_ripple():
    ny, R, U0 = 24, 4.0, 0.01
    H = ny - 2
    T = np.linspace(0.0, 10.0, 20001)
    yn, _, _ = _damped(T)
    ripple = 0.005 * np.sin(2 * np.pi * T / 0.05)
    Y = 0.5 * (ny - 1) + (yn + ripple) * (0.5 * H - R)
    _, _, idx = A.oscillation_extrema(T * H / U0, Y, ny, R, U0, prominence=0.05)
    _, _, clean = A.oscillation_extrema(T * H / U0, 0.5 * (ny - 1) + yn * (0.5 * H - R), ny, R, U0,
                                        prominence=0.05)
    assert len(idx) == len(clean)
    assert np.all(np.abs(T[idx] - T[clean]) < 0.05)
    # with a prominence below the ripple's, the ripple's own extrema appear
    _, _, fine = A.oscillation_extrema(T * H / U0, Y, ny, R, U0, prominence=0.001)
    assert len(fine) > 10 * len(clean)

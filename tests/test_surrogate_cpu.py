"""Host logic of the velocity surrogate (lbm_solver_3d_amd/surrogate.py):
the wall frame, the features and targets and their inverse, the free-space
model, and the channel draws."""
import math

import numpy as np
import pytest

from lbm_solver_3d_amd import surrogate as S


def test_mirrored_configurations_have_equal_features_and_mirrored_targets():
    nz, R, beta, B1 = 30, 4.0, -1.5, 1e-3
    X = np.array([10.0, 7.0, 8.25])                  # nearer the lower wall (z = 0.5)
    e = np.array([math.cos(0.3), 0.0, math.sin(0.3)])
    Xm = np.array([10.0, 7.0, (nz - 1.5) - (X[2] - 0.5)])  # mirror image near the upper wall
    em = np.array([e[0], 0.0, -e[2]])
    assert np.allclose(S.features(X, e, R, beta, nz), S.features(Xm, em, R, beta, nz))
    f = S.features(X, e, R, beta, nz)
    assert f[0] == pytest.approx((8.25 - 0.5) / R)
    assert f[1] == pytest.approx(math.sin(0.3)) and f[2] == pytest.approx(math.cos(0.3))
    assert f[3] == beta and f[4] == pytest.approx(R / (nz - 2))
    U, W = np.array([1e-4, 0.0, -2e-5]), np.array([0.0, 3e-6, 0.0])
    Um, Wm = np.array([1e-4, 0.0, 2e-5]), np.array([0.0, -3e-6, 0.0])   # reflected: z-velocity, x/y-rotation flip
    assert np.allclose(S.targets(X, U, W, R, B1, nz), S.targets(Xm, Um, Wm, R, B1, nz))


@pytest.mark.parametrize("z", [6.0, 22.0])
def test_to_lab_inverts_targets(z):
    nz, R, B1 = 30, 5.0, 1e-3
    X = np.array([3.0, 4.0, z])
    U, W = np.array([2e-4, 1e-5, -7e-5]), np.array([1e-6, -4e-6, 2e-6])
    y = S.targets(X, U, W, R, B1, nz)
    U2, W2 = S.to_lab(y, X, R, B1, nz)
    assert np.allclose(U2, U, rtol=1e-12, atol=0) and np.allclose(W2, W, rtol=1e-12, atol=0)


def test_free_space_model_is_blake_swimming():
    x = np.array([[2.0, math.sin(0.2), math.cos(0.2), 1.0, 0.2]])
    y = S.free_space(x)
    assert np.allclose(y[0], [math.cos(0.2), 0, math.sin(0.2), 0, 0, 0])


def test_channel_draws_are_valid_and_reproducible(lbm):
    a = S.channel_configs(40, 3)
    assert a == S.channel_configs(40, 3)
    for c in a:
        H, R = c["nz"], c["radius"]
        assert 24 <= H <= 64 and 4 <= R <= 8 and (c["nx"], c["ny"]) == (4 * H, 2 * H)
        assert c["b1"] == 1e-3 and c["walls_z"] == 1 and c["plane_xz"] == 1 and c["hmin"] == 1.0
        assert -5.0 <= c["beta"] <= 5.0 and abs(c["angle"]) < 45.0
        d, _ = S.wall_frame(np.array([0, 0, c["z0"]]), H)
        assert d - R >= 1.5 - 1e-9
        lbm.validate(**c)


class _Fixed(S.Surrogate):
    """The integrator with a given prediction in place of the networks."""

    def __init__(self, y):
        self.y = np.asarray(y, float)

    def predict(self, x):
        return np.tile(self.y, (len(x), 1)), np.zeros(len(x))


def test_advance_free_swimming_is_a_straight_line():
    """Free-space outputs (U along e, no rotation): every configuration moves
    U0 e per step, whatever side its nearest wall is on."""
    cfg = [dict(radius=5.0, beta=0.0, nz=40, b1=1e-3, hmin=1.0)] * 2
    th = 0.3
    X0 = np.array([[10.0, 3.0, 12.0], [10.0, 3.0, 27.0]])      # nearer the lower / the upper wall
    e0 = np.array([[math.cos(th), 0.0, math.sin(th)]] * 2)
    # wall-frame outputs of free swimming: U = U0 (cos, 0, s sin), s the side
    X, e = _Fixed([math.cos(th), 0, math.sin(th), 0, 0, 0]).advance(cfg[:1], X0[:1], e0[:1], 1000, 100)
    assert np.allclose(X[0], X0[0] + 1000 * (2e-3 / 3) * e0[0]) and np.allclose(e, e0)
    X, e = _Fixed([math.cos(th), 0, -math.sin(th), 0, 0, 0]).advance(cfg[1:], X0[1:], e0[1:], 1000, 100)
    assert np.allclose(X[0], X0[1] + 1000 * (2e-3 / 3) * e0[1]) and np.allclose(e, e0[1:])


def test_advance_rotation_and_hard_gap():
    """A wall-frame rotation rate w (Omega_y R / U0) turns e at w U0 / R per
    step about y (toward +x for the lower wall's frame); the centre stays at
    least R + hmin + 0.5 from the wall it swims into."""
    cfg = [dict(radius=4.0, beta=0.0, nz=30, b1=1e-3, hmin=1.0)]
    U0, R = 2e-3 / 3, 4.0
    X0 = np.array([[0.0, 0.0, 7.0]])
    e0 = np.array([[1.0, 0.0, 0.0]])
    w = 0.5
    X, e = _Fixed([0, 0, 0, 0, w, 0]).advance(cfg, X0, e0, 100, 1)
    ang = math.atan2(-e[0, 2], e[0, 0])   # Omega_y > 0 turns +x toward -z
    assert ang == pytest.approx(100 * w * U0 / R, rel=1e-4)
    X, e = _Fixed([0, 0, -5.0, 0, 0, 0]).advance(cfg, X0, e0, 1000, 10)   # straight into the lower wall
    assert X[0, 2] == pytest.approx(0.5 + R + 1.0)

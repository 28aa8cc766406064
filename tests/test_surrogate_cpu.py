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
        assert 16 <= H <= 32 and 3 <= R <= 6 and (c["nx"], c["ny"]) == (4 * H, 2 * H)
        d, _ = S.wall_frame(np.array([0, 0, c["z0"]]), H)
        assert d - R >= 1.5 - 1e-9
        lbm.validate(**c)

"""The learned velocity surrogate end to end on the GPU, at reduced size:
training data from batched lattice-Boltzmann runs, the 4-net ensemble
trained by the library's kernels, its held-out error against the model
without it (free-space swimming), and trajectories with LBM fallback against
full-LBM FP64 runs."""
import pytest

from lbm_solver_3d_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def report():
    return S.study(n_train_runs=256, n_test_runs=64, n_traj=64, n_ref=4, seed=0,
                   sample_at=tuple(range(3000, 5001, 250)), epochs=400, traj_warmup=3000, traj_steps=2000, every=100)


def test_surrogate_beats_free_space_model(report):
    h = report["held_out"]
    print(report)
    # the network captures what the walls do to the swimmer; the model without it cannot
    assert h["rms_mlp"] < 0.5 * h["rms_free_space"], h
    assert h["rms_mlp_speed"] < 0.5 * h["rms_free_space_speed"], h


def test_trajectories_against_fp64_lbm(report):
    v = report["vs_fp64"]
    t = report["trajectories"]
    assert t["count"] == 64 and 0 <= t["fallback"] <= 64
    # position error after 2000 steps, relative to the distance travelled
    assert v["mean_position_error"] < 0.25 * v["mean_travel"], v
    assert v["mean_position_error"] < v["mean_position_error_free_space"], v

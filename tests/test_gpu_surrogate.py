"""The learned velocity surrogate end to end on the GPU, at reduced channel
sizes (H in [24, 40]; the benchmark's surrogate workload uses H in
[24, 64]): training data from batched lattice-Boltzmann runs of slit
channels at B1 = 1e-3, the 4-net ensemble trained by the library's kernels,
its held-out error against the model without it (free-space swimming), 1000
seeded trajectories driven by the surrogate with the LBM run over the
windows where the ensemble disagrees, and their error against full-LBM FP64
runs of 16 of them."""
import pytest

from lbm_solver_3d_amd import surrogate as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def report():
    print("[surrogate] reduced study: 128 + 48 training / held-out runs, 1000 trajectories", flush=True)
    return S.study(n_train_runs=128, n_test_runs=48, n_traj=1000, n_ref=16, seed=0, H=(24, 40), warmup=3000,
                   window=3000, windows=4, epochs=600, traj_warmup=1500, traj_steps=3000, every=100,
                   traj_window=1000, fallback_warmup=1000, log=lambda m: print(m, flush=True))


def test_surrogate_beats_free_space_model(report):
    h = report["held_out"]
    print(report)
    # the network captures what the walls do to the swimmer; the model without it cannot
    assert h["rms_mlp"] < 0.5 * h["rms_free_space"], h
    assert h["rms_mlp_speed"] < 0.5 * h["rms_free_space_speed"], h


def test_trajectories_against_fp64_lbm(report):
    v = report["vs_fp64"]
    t = report["trajectories"]
    assert t["count"] == 1000 and t["windows"] == 3
    assert 0 <= t["fallback_windows"] <= 3000
    # position error after 3000 steps, relative to the distance travelled
    assert v["runs"] == 16
    assert v["mean_position_error"] < 0.25 * v["mean_travel"], v
    assert v["mean_position_error"] < v["mean_position_error_free_space"], v

"""The streaming-copy bandwidth probe the bench reports beside the HBM peak."""
import pytest

pytestmark = pytest.mark.gpu


def test_copy_bandwidth_is_plausible(lbm):
    # 1 GiB per buffer: well beyond the caches, so the copy streams from HBM
    gbs = lbm.copy_bandwidth(2 << 30, reps=5)
    print("copy GB/s", gbs)
    assert 1000.0 < gbs < 8000.0  # below the published 8 TB/s peak

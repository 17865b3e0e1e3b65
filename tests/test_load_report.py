"""Host logic of the per-wavefront load report (fcs.lib.wavefront_loads /
load_summary): the attempts each step-kernel wavefront makes follow from
the list lengths of the launches, since a launch strides its n cells over
the grid (cell i of the list goes to wavefront i mod grid)."""
import numpy as np

from fcs import lib


def brute(list_len, grid):
    loads = np.zeros(grid, np.int64)
    for n in list_len:
        for i in range(n):
            loads[i % grid] += 1
    return loads


def test_wavefront_loads_match_the_striding():
    rng = np.random.default_rng(0)
    for grid in (1, 7, 64, 1024):
        lens = rng.integers(0, 3 * grid + 5, 30)
        np.testing.assert_array_equal(lib.wavefront_loads(lens, grid), brute(lens, grid))


def test_load_summary():
    s = lib.load_summary([10, 4, 1], 4)
    # wavefronts: [3+1+1, 3+1, 2+1, 2+1] = [5, 4, 3, 3]
    assert s["attempts"] == 15 and s["launches"] == 3
    assert s["attempts_per_wavefront"] == {"min": 3, "mean": 3.75, "max": 5}
    assert s["tail_launches"] == 1 and s["tail_attempt_share"] == 1 / 15
    assert s["mean_busy_wavefronts_per_launch"] == (4 + 4 + 1) / 3

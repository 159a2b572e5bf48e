"""Request generation of the serving loop (CPU)."""
import numpy as np

from athad_cond_amd import serving as S


def test_poisson_arrivals_are_seeded_and_have_the_stated_rate():
    a = S.poisson_requests(400, rate_per_s=4.0, seed=12)
    b = S.poisson_requests(400, rate_per_s=4.0, seed=12)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 0.25) < 0.04            # mean 1/rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.2   # exponential: coefficient of variation 1
    grids = {r.grid for r in a}
    assert grids == set(S.GRID_CHOICES)
    assert all(r.steps == 50 and r.precision == "mixed" for r in a)

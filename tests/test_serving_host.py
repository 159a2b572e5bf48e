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


class FakeIntegration:
    """Stands in for a device context: work takes simulated time proportional to cells."""
    live = 0
    created = 0

    def __init__(self, grid, params, precision="fp64"):
        self.grid, self.precision = grid[0] if isinstance(grid, list) else grid, precision
        self.until = 0.0
        self.seed = None
        FakeIntegration.live += 1
        FakeIntegration.created += 1

    def init_state(self, seed):
        self.seed = seed

    def set_member_params(self, params, member=0):
        pass

    def run(self, n, graph=False):
        import time
        cells = self.grid[0] * self.grid[1] * self.grid[2]
        self.until = time.perf_counter() + n * cells * 2e-10

    def busy(self):
        import time
        return time.perf_counter() < self.until

    def field_stats(self, name, member=0):
        bad = self.precision == "mixed" and self.seed == 1003     # this request "blows up" in FP32
        return (float("nan"), 1.0) if bad else (float(self.seed), 2.0)

    def get(self, name, member=0):
        return np.zeros(self.grid)

    def close(self):
        FakeIntegration.live -= 1


def test_scheduler_admission_pooling_and_fallback(monkeypatch):
    monkeypatch.setattr(S, "Integration", FakeIntegration)
    monkeypatch.setattr(S, "footprint_bytes", lambda grid, prec: grid[0] * grid[1] * grid[2] * (4 if prec == "mixed" else 8))
    grids = [(11, 21, 41), (15, 31, 61)]
    reqs = S.poisson_requests(12, rate_per_s=200.0, seed=4, steps=8, grids=grids)
    budget = 3 * 15 * 31 * 61 * 4
    server = S.Server(hbm_budget=budget, max_slots=4, chunk=2)
    stats = server.run(reqs)
    assert stats["peak_hbm_bytes"] <= budget
    assert stats["contexts_reused"] > 0
    assert FakeIntegration.created < len(reqs) + 1          # pooled, not one context per request
    assert all(r.finished >= r.arrival for r in reqs)
    bad = [r for r in reqs if r.seed == 1003]
    assert bad and bad[0].fallback and bad[0].precision == "fp64" and stats["fallbacks"] == 1
    assert bad[0].result["t"] == (1003.0, 2.0)
    server.close()
    assert FakeIntegration.live == 0

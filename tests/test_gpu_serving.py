"""Serving loop on the GPU: requests that arrive while others run are admitted and overlap on
the device, admission respects the HBM budget, and each request's result is the same alone and
in the mix (bit for bit: deterministic kernels, separate buffers)."""
import numpy as np
import pytest

from athad_cond_amd import serving as S
from athad_cond_amd.integration import footprint_bytes

pytestmark = pytest.mark.gpu

GRIDS = [(11, 21, 41), (21, 91, 181), (15, 31, 61)]


def test_same_result_alone_and_in_mix():
    print("case serving alone vs mix", flush=True)
    reqs = S.poisson_requests(9, rate_per_s=500.0, seed=3, steps=10, grids=GRIDS)
    server = S.Server(max_slots=4, chunk=2, keep_fields=True)
    stats = server.run(reqs)
    print(stats, flush=True)
    assert stats["peak_active"] > 1
    assert stats["contexts_reused"] > 0          # later requests ran in pooled contexts
    for r in reqs[:4]:
        alone = S.Request(r.rid, 0.0, r.grid, r.state, r.steps, "mixed", r.seed)
        one = S.Server(max_slots=1, keep_fields=True)
        one.run([alone])
        one.close()
        for n, v in alone.result["fields"].items():
            assert np.array_equal(v, r.result["fields"][n]), (r.rid, n)
            assert alone.result[n] == r.result[n], (r.rid, n)   # device digests, bit for bit
    server.close()


def test_admission_respects_budget():
    print("case serving admission", flush=True)
    reqs = S.poisson_requests(5, rate_per_s=1000.0, seed=8, steps=6, grids=[(21, 91, 181)])
    one = footprint_bytes((21, 91, 181), "mixed")
    server = S.Server(hbm_budget=2 * one + 1, max_slots=8, chunk=2)
    stats = server.run(reqs)
    server.close()
    print(stats, flush=True)
    assert stats["peak_active"] == 2
    assert stats["peak_hbm_bytes"] <= 2 * one + 1
    assert all(r.finished > 0 for r in reqs)
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0

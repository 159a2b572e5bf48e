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


def test_c5_grids_alone_and_in_mix():
    """The benchmark's serving grids, the largest two included (41 x 181 x 361 and
    61 x 361 x 721, a few iterations each): requests admitted together overlap on the device
    and each result equals the same request served alone, bit for bit."""
    print("case serving C5 grids alone vs mix", flush=True)
    grids = [(41, 181, 361), (61, 361, 721), (21, 91, 181)]
    reqs = S.poisson_requests(6, rate_per_s=200.0, seed=17, steps=4, grids=grids)
    got = {tuple(r.grid) for r in reqs}
    assert (61, 361, 721) in got and (41, 181, 361) in got, got
    server = S.Server(max_slots=3, chunk=2)
    stats = server.run(reqs)
    server.close()
    print(stats, flush=True)
    assert stats["peak_active"] > 1
    for r in reqs:
        if tuple(r.grid) == (21, 91, 181):
            continue
        alone = S.Request(r.rid, 0.0, r.grid, r.state, r.steps, "mixed", r.seed)
        one = S.Server(max_slots=1)
        one.run([alone])
        one.close()
        for n in alone.result:
            if n != "fields":
                assert alone.result[n] == r.result[n], (r.rid, r.grid, n)

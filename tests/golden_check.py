"""Comparison of a state against the reference's golden vectors (tests/golden/*.npz): a fixed
sample of cells per field plus every field's sum and L2 norm, as written by
tests/golden/make_ref_vectors.py from the reference's own compiled code."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GRIDS = os.path.join(GOLDEN, "ref_grids.npz")


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.sqrt(np.sum(b * b))
    num = np.sqrt(np.sum((a - b) ** 2))
    return float(num / den) if den > 0 else float(num)


def pack_errors(get, gold, prefix, names, index_key="index"):
    """Worst of the sample relative L2, the sum's and the L2 norm's relative deviation, per
    field; get(name) returns the whole field."""
    idx = gold[index_key]
    errs = {}
    for name in names:
        if f"{prefix}/{name}/sample" not in gold:
            continue
        a = np.asarray(get(name), dtype=np.float64).reshape(-1)
        e = rel(a[idx], gold[f"{prefix}/{name}/sample"])
        s_ref, l_ref = float(gold[f"{prefix}/{name}/sum"]), float(gold[f"{prefix}/{name}/l2"])
        s, l2 = float(np.sum(a)), float(np.sqrt(np.sum(a * a)))
        e = max(e, abs(s - s_ref) / max(abs(s_ref), l_ref, 1e-300), abs(l2 - l_ref) / max(l_ref, 1e-300))
        errs[name] = e
    return errs


def grid_run_tag(grid, seed, steps):
    return "%dx%dx%d/seed%d/steps%d" % (tuple(grid) + (seed, steps))


def grid_run_errors(get, gold, grid, seed, steps, names):
    tag = grid_run_tag(grid, seed, steps)
    return pack_errors(get, gold, tag, names, index_key=f"{tag}/index")

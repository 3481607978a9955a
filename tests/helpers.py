"""Comparison helpers shared by the parity tests."""
import numpy as np


def rownorm_rel(a, b):
    """Row-wise relative error ||a_r - b_r|| / ||b_r|| (a deviational
    temperature crosses zero inside the domain, so per-cell relative error is
    not meaningful)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for r in range(b.shape[0]):
        nb = np.linalg.norm(b[r])
        out.append(np.linalg.norm(a[r] - b[r]) / nb if nb > 0 else np.linalg.norm(a[r]))
    return np.array(out)


def first_mismatch(a, b):
    idx = np.nonzero(np.asarray(a) != np.asarray(b))[0]
    return int(idx[0]) if idx.size else -1


def assert_counts_equal(cg, co):
    for name in ("nscat", "nbdry", "nstep", "fate"):
        g, o = getattr(cg, name), getattr(co, name)
        k = first_mismatch(g, o)
        assert k < 0, f"{name} differs first at bundle {k}: gpu {g[k]} oracle {o[k]}"

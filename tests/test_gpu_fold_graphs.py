"""Fold steps replayed as HIP graphs (RCWA_FOLD_GRAPHS=1, FP64 form) against
the same steps launched directly (the default): bit for bit, on chunks
whose structures have different layer counts (so that the number of
products still folding changes and steps are captured anew), over several
solves on one context with the layer stage's own work running between the
replays, and at the 23 x 23-harmonic size where a step holds the
block-recursive inverse."""
import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def _chunk(seed, h, grid, layers):
    rng = np.random.default_rng(seed)
    return [W.random_small(rng, hx=h, hy=h, L=L, grid=grid) for L in layers]


def _same(a, b):
    return all(np.array_equal(np.asarray(x), np.asarray(y)) for x, y in zip(a, b))


@pytest.mark.parametrize("h,grid,layers", [(5, 32, (8, 6, 3, 8)), (11, 64, (4, 2))])
def test_fold_graphs_equal_direct_launches(monkeypatch, h, grid, layers):
    probs = _chunk(100 + h, h, grid, layers)
    ctx = G.Context(0)
    try:
        monkeypatch.setenv("RCWA_FOLD_GRAPHS", "1")
        first = ctx.solve(probs)
        again = ctx.solve(probs)
        monkeypatch.delenv("RCWA_FOLD_GRAPHS")
        direct = ctx.solve(probs)
        monkeypatch.setenv("RCWA_FOLD_GRAPHS", "1")
        third = ctx.solve(probs)
        monkeypatch.delenv("RCWA_FOLD_GRAPHS")
    finally:
        ctx.close()
    for i in range(len(probs)):
        assert _same(first[i], direct[i]), i
        assert _same(again[i], direct[i]), i
        assert _same(third[i], direct[i]), i

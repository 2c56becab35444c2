"""A single small layer batch formed as staggered groups on their own streams
(each group's stage starts behind the previous group's Hessenberg
reduction; two groups by default, three with RCWA_LAYER_GROUPS=3) against
the same batch formed in one piece (RCWA_LAYER_HALVES=0), on the same
chunks: one structure of five layers at 15 x 15 harmonics (n = 450, where
the groups are the default; 3 + 2, or 1 + 2 + 2), three structures of three
layers (nine layer matrices, 5 + 4: layer 1's three straddle the halves, or
3 + 3 + 3), and the groups forced on at 11 x 11 harmonics
(n = 242).  Each matrix goes through the same kernels either way, so the
two agree to rounding; the lossless cases also check R + T = 1."""
import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def _problems(seed, h, count, L, lossless):
    rng = np.random.default_rng(seed)
    # a grid of side at least 4h: the Fourier coefficients up to order 2h
    # without aliasing
    return [W.random_small(rng, hx=h, hy=h, L=L, grid=64, lossless=lossless)
            for _ in range(count)]


@pytest.mark.parametrize("h,count,L,lossless,force,groups", [
    (7, 1, 5, False, False, 2),
    (7, 1, 5, False, False, 3),
    (7, 3, 3, True, False, 2),
    (7, 3, 3, True, False, 3),
    (5, 2, 4, False, True, 2),
])
def test_layer_halves_match_one_batch(ctx, monkeypatch, h, count, L, lossless, force, groups):
    probs = _problems(2000 + 10 * h + count + L, h, count, L, lossless)
    if force:
        monkeypatch.setenv("RCWA_LAYER_HALVES", "1")
    else:
        monkeypatch.delenv("RCWA_LAYER_HALVES", raising=False)
    if groups == 3:
        monkeypatch.setenv("RCWA_LAYER_GROUPS", "3")
    else:
        monkeypatch.delenv("RCWA_LAYER_GROUPS", raising=False)
    two = ctx.solve(probs)
    monkeypatch.delenv("RCWA_LAYER_GROUPS", raising=False)
    monkeypatch.setenv("RCWA_LAYER_HALVES", "0")
    one = ctx.solve(probs)
    monkeypatch.delenv("RCWA_LAYER_HALVES")
    for a, b in zip(two, one):
        d = max(float(np.abs(np.asarray(x) - np.asarray(y)).max()) for x, y in zip(a, b))
        print("%d groups against one batch (h = %d, %d x %d layers): max diff %.2e"
              % (groups, h, count, L, d))
        assert d < 1e-10, d
        if lossless:
            e = float(np.abs(a[2].sum(axis=1) - 1.0).max())
            print("lossless: R + T - 1 = %.2e" % e)
            assert e < 1e-10, e


def test_layer_halves_timing_counts_both(ctx, monkeypatch):
    """With the eigensolver timers on, every group is counted (four layer
    matrices: groups of 2 and 2)."""
    probs = _problems(2100, 7, 1, 4, False)
    monkeypatch.delenv("RCWA_LAYER_HALVES", raising=False)
    ctx.set_timing(True)
    try:
        ctx.reset_stats()
        ctx.solve(probs)
        st = ctx.stats()
    finally:
        ctx.set_timing(False)
    print("eigensolver launches %d, matrices %d" % (st["n_eig_launches"], st["n_eig_matrices"]))
    assert st["n_eig_launches"] == 2
    assert st["n_eig_matrices"] == 4

"""The fold's inverse of I - X11 P22 for n > 512 in batches of at most 64
(2x2-block recursion on the Schur complement, Gauss-Jordan inside the blocks
of at most 512 rows) against the Gauss-Jordan inverse for every step
(RCWA_FOLD_INVERSE=gj), on the same chunks: at the benchmark's 21 x 21
harmonics (n = 882, one level) and at 23 x 23 (n = 1058), alone and three
at a time.  The fold's refined solves make the two agree to rounding."""
import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def _problems(seed, h, count, lossless):
    rng = np.random.default_rng(seed)
    # a 64 x 64 grid: the Fourier coefficients up to order 2h without aliasing
    return [W.random_small(rng, hx=h, hy=h, L=3, grid=64, lossless=lossless)
            for _ in range(count)]


@pytest.mark.parametrize("h,count,lossless", [(10, 3, False), (11, 1, False), (11, 3, True)])
def test_fold_block_inverse_matches_gauss_jordan(ctx, monkeypatch, h, count, lossless):
    probs = _problems(1000 + 10 * h + count + int(lossless), h, count, lossless)
    monkeypatch.delenv("RCWA_FOLD_INVERSE", raising=False)
    blk = ctx.solve(probs)
    monkeypatch.setenv("RCWA_FOLD_INVERSE", "gj")
    gj = ctx.solve(probs)
    monkeypatch.delenv("RCWA_FOLD_INVERSE")
    for b, g in zip(blk, gj):
        d = max(float(np.abs(np.asarray(x) - np.asarray(y)).max()) for x, y in zip(b, g))
        print("block-inverse fold against Gauss-Jordan fold: max diff %.2e" % d)
        assert d < 1e-10, d
        if lossless:
            e = float(np.abs(b[2].sum(axis=1) - 1.0).max())
            print("lossless: R + T - 1 = %.2e" % e)
            assert e < 1e-10, e

"""The fold's inverse of I - X11 P22 for n > 1024 in batches of at most 8
(2x2-block recursion on the Schur complement, Gauss-Jordan inside the blocks
of at most 1024 rows) against the Gauss-Jordan inverse the fold takes for
larger batches, on the same problem: a structure solved alone (batch 1,
block inverse) and in a chunk of nine (Gauss-Jordan).  The fold's refined
solves make the two agree to rounding."""
import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def _problems(seed, count, lossless):
    rng = np.random.default_rng(seed)
    # 23 x 23 harmonics: n = 2 * 529 = 1058 > 1024
    # a 64 x 64 grid: the Fourier coefficients up to order 22 without aliasing
    return [W.random_small(rng, hx=11, hy=11, L=3, grid=64, lossless=lossless)
            for _ in range(count)]


@pytest.mark.parametrize("lossless", [False, True])
def test_fold_block_inverse_matches_gauss_jordan(ctx, lossless):
    probs = _problems(1058 + int(lossless), 9, lossless)
    alone = ctx.solve(probs[:1])[0]
    chunk = ctx.solve(probs)[0]
    for a, c in zip(alone, chunk):
        d = float(np.abs(np.asarray(a) - np.asarray(c)).max())
        print("block-inverse fold against Gauss-Jordan fold: max diff %.2e" % d)
        assert d < 1e-10, d
    if lossless:
        e = float(np.abs(alone[2].sum(axis=1) - 1.0).max())
        print("lossless, batch 1: R + T - 1 = %.2e" % e)
        assert e < 1e-10, e

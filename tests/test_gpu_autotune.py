"""Tile-shape autotuning of the complex GEMMs (rcwa_ctx_set_autotune): every
tile shape gives the same bits, the tuner records the shape classes it meets,
and solves and served requests with tuning on equal those without it."""
import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu

NSHAPES = 5


def rnd(rng, m, n):
    return rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))


@pytest.mark.parametrize("m,n,k,b", [(882, 882, 882, 3), (242, 2, 242, 5), (70, 33, 100, 2),
                                     (98, 98, 16, 7)])
def test_every_tile_shape_gives_the_same_bits(ctx, m, n, k, b):
    rng = np.random.default_rng(m + n + k)
    A = [rnd(rng, m, k) for _ in range(b)]
    B = [rnd(rng, k, n) for _ in range(b)]
    try:
        G.zgemm_force_shape(0)
        ref = ctx.zgemm(A, B)
        for v in range(1, NSHAPES):
            G.zgemm_force_shape(v)
            C = ctx.zgemm(A, B)
            for c, r in zip(C, ref):
                assert np.array_equal(c, r), "tile shape %d differs" % v
    finally:
        G.zgemm_force_shape(-1)
    err = max(np.abs(c - a @ bb).max() / np.abs(a @ bb).max() for c, a, bb in zip(ref, A, B))
    assert err < 1e-13


def test_no_such_shape_is_refused():
    with pytest.raises(G.RcwaError):
        G.zgemm_force_shape(NSHAPES)


def test_tuned_solve_equals_untuned(ctx):
    rng = np.random.default_rng(5)
    probs = [W.random_small(rng, hx=3, hy=3, L=3, grid=32) for _ in range(4)]
    plain = ctx.solve(probs)
    t = G.Context(0)
    try:
        t.set_autotune(True)
        tuned = t.solve(probs)
        again = t.solve(probs)  # the tuned shapes, no tuning
    finally:
        t.close()
    for a, b, c in zip(plain, tuned, again):
        for x, y, z in zip(a[:3], b[:3], c[:3]):
            assert np.array_equal(x, y) and np.array_equal(x, z)
    ent = G.autotune_entries()
    assert ent, "the tuner recorded no shape class"
    for M, N, K, bclass, dev, shape in ent:
        assert M > 0 and N > 0 and K >= 0 and 0 <= shape < NSHAPES


def test_server_with_autotune_matches_per_call(ctx):
    rng = np.random.default_rng(8)
    probs = [W.random_small(rng, hx=int(h), hy=int(h), L=int(L), grid=16)
             for h, L in zip(rng.integers(1, 4, 10), rng.integers(1, 5, 10))]
    with G.Server(0, hbm_budget=8 << 30, n_workers=2, autotune=True) as srv:
        res = [srv.wait(srv.submit(p)) for p in probs]
    for p, r in zip(probs, res):
        alone = ctx.solve([p])[0]
        assert np.array_equal(r[0], alone[0]) and np.array_equal(r[1], alone[1])

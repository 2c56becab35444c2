"""rcwa_star_product_batched_noalloc, the allocation-free form of the
batched, HBM-resident Redheffer product (no counterpart in the reference;
for callers that cannot allocate device memory): on two calls in a row, the
second with a smaller input, its results equal those of
rcwa_star_product_batched (same kernels, work buffers in the library's
fixed scratch array instead of the context's arena) bit for bit, and sizes
beyond its limits are refused with RCWA_E_ARG."""
import numpy as np
import pytest

import rcwa_amd as G

pytestmark = pytest.mark.gpu


def _colmajor(mats):
    return np.ascontiguousarray(np.stack([m.T for m in mats]))


def _run(ctx, n, b, seed):
    rng = np.random.default_rng(seed)
    blocks = [0.3 * (rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n)))
              for _ in range(8)]
    dev = [G.DeviceBuffer(b * n * n * 16) for _ in range(16)]
    try:
        for k in range(8):
            dev[k].upload(_colmajor(blocks[k]))
        A, B = [x.ptr for x in dev[0:4]], [x.ptr for x in dev[4:8]]
        S1, S2 = [x.ptr for x in dev[8:12]], [x.ptr for x in dev[12:16]]
        ctx.star_batched(n, b, A, B, S1)
        ctx.star_batched(n, b, A, B, S2, noalloc=True)
        r1 = [dev[8 + k].download((b, n * n)) for k in range(4)]
        r2 = [dev[12 + k].download((b, n * n)) for k in range(4)]
    finally:
        for x in dev:
            x.free()
    for k in range(4):
        assert np.all(np.isfinite(r2[k]))
        assert np.array_equal(r1[k], r2[k]), (n, b, k, np.abs(r1[k] - r2[k]).max())


def test_noalloc_star_product_equals_allocating_form():
    with G.Context(0) as ctx:
        _run(ctx, 98, 8, 1)     # 98^2 * 8 complex values per slot
        _run(ctx, 40, 3, 2)     # then a smaller input on the same scratch
        with pytest.raises(G.RcwaError) as e:
            d = G.DeviceBuffer(16)
            try:
                ctx.star_batched(2048, 1, [d.ptr] * 4, [d.ptr] * 4, [d.ptr] * 4, noalloc=True)
            finally:
                d.free()
        assert e.value.code == 1

"""Device rasterisation (the step before ConvMat, main.cpp:48-91 in the
reference) against the host generators of the workloads: cell-centre
samples, same inside tests, so the grids must be equal bit for bit; and a
solve from a rasterised device grid against the same solve from host grids."""
import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def test_rasterize_c1_hole(ctx):
    Lx, Ly, bg, shapes = W.c1_shapes()
    g = ctx.rasterize(96, 80, Lx, Ly, bg, shapes)
    ref = W.circle_hole_grid(96, 80, Lx, Ly, 150.0, 1.0, 15.0 + 0.1j)
    assert np.array_equal(g.download()[0], ref)


def test_rasterize_c2_line(ctx):
    Lx, Ly, bg, shapes = W.c2_shapes(14)
    g = ctx.rasterize(128, 16, Lx, Ly, bg, shapes)
    ref = W.c2_sweep(indices=[14], grid=128, h=1)[0].er[:, :16, :]
    assert np.array_equal(g.download(), ref)


def test_rasterize_c4_layers(ctx):
    rng = np.random.default_rng(44)
    shapes, bg, ref = [], [], []
    for layer in range(6):
        grid, eb = W.random_shapes_layer(rng, 64, 500.0, 500.0, shapes, layer)
        bg.append(eb)
        ref.append(grid)
    g = ctx.rasterize(64, 64, 500.0, 500.0, bg, shapes)
    assert np.array_equal(g.download(), np.stack(ref))


def test_rasterize_c5_holes(ctx):
    Lx, Ly, bg, shapes = W.c5_shapes()
    # shapes listed layer-major out of order: the entry point sorts by layer
    # and keeps each layer's order
    shapes = shapes[5:] + shapes[:5]
    g = ctx.rasterize(128, 128, Lx, Ly, bg, shapes)
    assert np.array_equal(g.download(), W.c5_large(grid=128).er)


def test_rasterize_rejects_bad_shapes(ctx):
    with pytest.raises(G.RcwaError):
        ctx.rasterize(8, 8, 1.0, 1.0, [1.0], [(1, G.SHAPE_CIRCLE, 0.5, 0.5, 0.2, 0.0, 2.0)])
    with pytest.raises(G.RcwaError):
        ctx.rasterize(8, 8, 1.0, 1.0, [1.0], [(0, 7, 0.5, 0.5, 0.2, 0.0, 2.0)])


def test_solve_from_rasterized_grid(ctx):
    """The line grating at one wavelength from host grids and from device
    grids rasterised from its shape list: the same grids, so the same R, T."""
    p = W.c2_sweep(indices=[70], grid=64, h=3)[0]
    host = ctx.solve([p], fields=True)[0]
    Lx, Ly, bg, shapes = W.c2_shapes(70)
    er = ctx.rasterize(64, 64, Lx, Ly, bg, shapes)
    ur = ctx.rasterize(64, 64, Lx, Ly, [1.0] * len(bg), [])
    dev = ctx.solve([p], device_grids=[(er.ptr, ur.ptr)], grid_shape=(64, 64), fields=True)[0]
    for a, b in zip(host, dev):
        assert np.array_equal(a, b)
    er.free()
    ur.free()

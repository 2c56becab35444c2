"""Full-size parity against the committed golden fixtures (tests/golden/,
written once by tests/golden/make_fixtures.py from the CPU FP64 oracle):

  - the benchmark's structure (21x21 harmonics, n = 882, 20 layers, 512^2
    grids) at four wavelengths of its 250-800 nm sweep, 288.69 nm -- the
    substrate cutoff of the (+1, 0) order, |k_z| ~ 0.04, the sweep's
    worst-conditioned point -- among them: per-order R and T within 1e-9 of
    the oracle for TE and TM (rcwa.cpp:564-585), in the FP64 form; the
    split-precision form within its 1e-6 target of the same oracle values;
  - the first two layers of the 41x41-harmonic hole-array stack (n = 3362,
    1024^2 grids) within 1e-9, and the whole 50-layer stack made lossless
    conserving energy, R + T = 1 to 1e-10."""
import os

import numpy as np
import pytest

import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C2_INDICES = (14, 60, 120, 199)


def _fixture(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        pytest.skip("fixture %s not generated (tests/golden/make_fixtures.py)" % name)
    return np.load(path, allow_pickle=False)


def _err(out, fx):
    R, T, tot = out[:3]
    return max(np.abs(R - fx["R"]).max(), np.abs(T - fx["T"]).max(),
               np.abs(tot - fx["tot"]).max())


@pytest.fixture(scope="module")
def c2_fixtures():
    fx = [_fixture("c2_lam%d.npz" % i) for i in C2_INDICES]
    probs = W.c2_sweep(indices=list(C2_INDICES))
    for p, f in zip(probs, fx):
        assert abs(p.lam0 - float(f["lam0"])) < 1e-12 and int(f["hx"]) == 10
    return probs, fx


def test_c2_full_size_vs_oracle_fixtures(ctx, c2_fixtures):
    probs, fx = c2_fixtures
    out = ctx.solve(probs)
    for i, o, f in zip(C2_INDICES, out, fx):
        e = _err(o, f)
        print("lambda index %d (%.2f nm): max |dR|, |dT| vs oracle %.2e" % (i, f["lam0"], e))
        assert e < 1e-9


def test_c2_split_form_vs_oracle_fixtures(ctx, c2_fixtures):
    probs, fx = c2_fixtures
    out = ctx.solve(probs, precision=G.RCWA_SPLIT)
    for i, o, f in zip(C2_INDICES, out, fx):
        e = _err(o, f)
        print("split form, lambda index %d: max |dR|, |dT| vs oracle %.2e" % (i, e))
        assert e < 1e-6


def _c5_solve(ctx, nlayer, lossless=False):
    """The first nlayer layers of the hole-array stack, grids rasterised on
    the device from the shape list (bit-equal to workloads.c5_large's host
    grids, which the fixture was computed from)."""
    Lx, Ly, bg, shapes = W.c5_shapes()
    bg = [complex(b.real) if lossless else b for b in bg[:nlayer]]
    shapes = [s for s in shapes if s[0] < nlayer]
    full = W.c5_large(grid=8)  # the structure's scalars; its grids come from the shapes
    p = W.Problem(full.er[:nlayer], full.ur[:nlayer], full.thick[:nlayer], full.Lx, full.Ly,
                  full.lam0, full.theta, full.phi, full.pols, full.hx, full.hy, full.er_ref,
                  full.ur_ref, full.er_trn, full.ur_trn)
    er = ctx.rasterize(1024, 1024, Lx, Ly, bg, shapes)
    ur = ctx.rasterize(1024, 1024, Lx, Ly, [1.0] * nlayer, [])
    try:
        return ctx.solve([p], device_grids=[(er.ptr, ur.ptr)], grid_shape=(1024, 1024))[0]
    finally:
        er.free()
        ur.free()


def test_c5_two_layers_vs_oracle_fixture(ctx):
    fx = _fixture("c5_2layers.npz")
    out = _c5_solve(ctx, 2)
    e = _err(out, fx)
    print("C5 first two layers (n = 3362): max |dR|, |dT| vs oracle %.2e" % e)
    assert e < 1e-10


def test_c5_full_stack_lossless_energy(ctx):
    """The 50-layer stack with eps = 4.0 + 0.05i replaced by 4.0: R + T = 1."""
    R, T, tot = _c5_solve(ctx, 50, lossless=True)
    dev = abs(tot.sum() - 1.0)
    print("C5 lossless, 50 layers: R = %.6f, T = %.6f, R + T - 1 = %.2e"
          % (tot[0, 0], tot[0, 1], dev))
    assert dev < 1e-10

"""Sweep plans (rcwa_sweep_*): the structure as material masks whose
convolution matrices are formed once, each point mixing them with its
material values.  Checked against the batched solve on the rasterised grids
(the same solve up to the rounding of the convolution matrices) and against
the FP64 oracle (per-order R/T within 1e-9)."""
import math

import numpy as np
import pytest

import oracle_py as O
import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-9        # against the oracle
TOL_DIRECT = 1e-11  # against the direct path on the same grids


def _max_diff(a, b):
    return max(np.abs(a[0] - b[0]).max(), np.abs(a[1] - b[1]).max(), np.abs(a[2] - b[2]).max())


def _problem(material, table, mu_table, thick, Lx, Ly, src, h, er_trn=2.25):
    er = np.asarray(table, np.complex128)[material]
    ur = (np.ones_like(er) if mu_table is None
          else np.asarray(mu_table, np.complex128)[material])
    lam, th, ph = src[0][:3]
    return W.Problem(er, ur, thick, Lx, Ly, lam, th, ph, [(s[3], s[4]) for s in src], h, h,
                     1.0, 1.0, er_trn, 1.0)


def test_sweep_line_grating_vs_direct_and_oracle(ctx):
    """The benchmark's line grating (20 slices, 5 dispersive materials,
    theta = 65 deg, TE and TM) at 64x64 grids and 7x7 harmonics, at
    wavelengths across the sweep: every point has its own eps row."""
    material, thick, lams, eps = W.c2_structure(grid=64)
    idx = [0, 14, 70, 199]
    th = math.radians(65.0)
    srcs = [[(lams[i], th, 0.0, 1.0, 0.0), (lams[i], th, 0.0, 0.0, 1.0)] for i in idx]
    with ctx.sweep(material, thick, 120.0, 120.0, 6, 3, 3) as sw:
        out = sw.run(srcs, eps[idx])
    direct = ctx.solve(W.c2_sweep(indices=idx, grid=64, h=3))
    for i, o, d in zip(idx, out, direct):
        assert _max_diff(o, d) < TOL_DIRECT, (i, _max_diff(o, d))
    for i, o in zip(idx[:2], out[:2]):
        p = W.c2_sweep(indices=[i], grid=64, h=3)[0]
        assert _max_diff(o, O.solve(p)) < TOL


@pytest.mark.parametrize("with_mu", [False, True])
def test_sweep_angle_shared_inverse(ctx, with_mu):
    """An angle sweep of non-dispersive media: every point has the same eps
    (and mu) row, so [[eps]]^-1 (and [[mu]]^-1) are formed once per layer and
    shared."""
    rng = np.random.default_rng(5)
    L, grid, h = 3, 16, 2
    material = rng.integers(0, 3, size=(L, grid, grid)).astype(np.uint8)
    table = [1.0, 6.0 - 0.4j, 2.5]
    mu_table = [1.0, 1.3, 0.9 - 0.1j] if with_mu else None
    thick = rng.uniform(50, 150, L)
    Lx, Ly = 420.0, 380.0
    srcs = [[(550.0, math.radians(a), math.radians(15.0), 1.0, 0.0),
             (550.0, math.radians(a), math.radians(15.0), 0.0, 1.0)] for a in (0, 20, 40, 60, 75)]
    eps = np.tile(np.asarray(table, np.complex128), (len(srcs), 1))
    mu = None if mu_table is None else np.tile(np.asarray(mu_table, np.complex128), (len(srcs), 1))
    with ctx.sweep(material, thick, Lx, Ly, 3, h, h) as sw:
        out = sw.run(srcs, eps, mu)
    probs = [_problem(material, table, mu_table, thick, Lx, Ly, s, h) for s in srcs]
    direct = ctx.solve(probs)
    for o, d, p in zip(out, direct, probs):
        assert _max_diff(o, d) < TOL_DIRECT
        assert _max_diff(o, O.solve(p)) < TOL


def test_sweep_dispersive_mu_and_fields(ctx):
    """Material values that change per point, mu != 1, per-order complex
    amplitudes, and a point whose row repeats an earlier one (shared
    inverses next to unshared ones in one chunk)."""
    rng = np.random.default_rng(8)
    L, grid, h = 2, 16, 2
    material = rng.integers(0, 4, size=(L, grid, grid)).astype(np.uint8)
    thick = rng.uniform(60, 140, L)
    Lx, Ly = 500.0, 450.0
    npt = 4
    eps = 1.0 + rng.uniform(1, 10, (npt, 4)) - 1j * rng.uniform(0, 1, (npt, 4))
    mu = 1.0 + rng.uniform(0, 0.5, (npt, 4)) + 0j
    eps[3], mu[3] = eps[1], mu[1]
    lams = [480.0, 520.0, 610.0, 700.0]
    srcs = [[(lam, math.radians(30.0), 0.0, 0.6, 0.8)] for lam in lams]
    with ctx.sweep(material, thick, Lx, Ly, 4, h, h) as sw:
        out = sw.run(srcs, eps, mu, fields=True)
    probs = [_problem(material, eps[p], mu[p], thick, Lx, Ly, srcs[p], h) for p in range(npt)]
    direct = ctx.solve(probs, fields=True)
    for o, d, p in zip(out, direct, probs):
        assert _max_diff(o, d) < TOL_DIRECT
        assert np.abs(o[3] - d[3]).max() < TOL_DIRECT
        assert np.abs(o[4] - d[4]).max() < TOL_DIRECT
        assert _max_diff(o, O.solve(p)) < TOL


def test_sweep_chunks_of_equal_size(ctx):
    """Seven points with at most three per chunk run as chunks of 3, 2 and 2:
    every matrix of the batched solve is processed independently, so the
    results equal those of one chunk bit for bit."""
    material, thick, lams, eps = W.c2_structure(grid=32, nslice=4)
    idx = [3, 20, 41, 77, 120, 150, 190]
    srcs = [[(lams[i], math.radians(40.0), 0.0, 1.0, 0.0)] for i in idx]
    with ctx.sweep(material, thick, 120.0, 120.0, 6, 2, 2) as sw:
        whole = sw.run(srcs, eps[idx])
        ctx.set_chunk(3)
        try:
            parts = sw.run(srcs, eps[idx])
        finally:
            ctx.set_chunk(0)
    for a, b in zip(whole, parts):
        assert _max_diff(a, b) == 0.0


def test_sweep_rejects_bad_material(ctx):
    material = np.zeros((1, 8, 8), np.uint8)
    material[0, 3, 3] = 2
    with pytest.raises(G.RcwaError) as e:
        ctx.sweep(material, [100.0], 300.0, 300.0, 2, 1, 1)
    assert e.value.code == 1

"""End-to-end parity of the HIP solve path with the CPU FP64 oracle.

Tolerances (from the FP64 contract): per-order and total R/T within 1e-9
absolute of the oracle; R + T = 1 within 1e-10 for lossless stacks."""
import math

import numpy as np
import pytest

import np_rcwa as NR
import oracle_py as O
import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-9


def _cmp(gpu, ref):
    R, T, tot = gpu
    Ro, To, toto = ref
    return (np.abs(R - Ro).max(), np.abs(T - To).max(), np.abs(tot - toto).max())


def test_layer_smatrix_vs_oracle(ctx):
    rng = np.random.default_rng(3)
    p = W.random_small(rng, hx=2, hy=2, L=2, grid=16)
    k, Kx, Ky, Kzr, Kzt = O.wave_vectors(p.Lx, p.Ly, p.lam0, p.theta, p.phi, p.hx, p.hy,
                                         p.er_ref, p.ur_ref, p.er_trn, p.ur_trn)
    ercs = [O.convmat(p.er[l], p.hx, p.hy) for l in range(2)]
    urcs = [O.convmat(p.ur[l], p.hx, p.hy) for l in range(2)]
    k0 = 2 * np.pi / p.lam0
    S11, S12 = ctx.layer_smatrix(ercs, urcs, Kx, Ky, k0 * p.thick)
    for l in range(2):
        r11, r12 = O.smatrix_layer(ercs[l], urcs[l], Kx, Ky, k0, p.thick[l])
        assert np.abs(S11[l] - r11).max() < 1e-10
        assert np.abs(S12[l] - r12).max() < 1e-10


def test_star_product_vs_oracle(ctx):
    rng = np.random.default_rng(5)
    n = 18
    A = tuple(0.3 * (rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n)))
              for _ in range(4))
    B = tuple(0.3 * (rng.standard_normal((2, n, n)) + 1j * rng.standard_normal((2, n, n)))
              for _ in range(4))
    S = ctx.star(A, B)
    for b in range(2):
        ref = O.redheffer([x[b] for x in A], [x[b] for x in B])
        for k in range(4):
            assert np.abs(S[k][b] - ref[k]).max() < 1e-12


@pytest.mark.parametrize("seed,hx,hy,L,lossless", [(0, 1, 1, 1, True), (1, 2, 1, 3, False),
                                                   (2, 2, 2, 2, True), (3, 3, 2, 4, False),
                                                   (4, 0, 3, 2, False)])
def test_solve_small_vs_oracle(ctx, seed, hx, hy, L, lossless):
    rng = np.random.default_rng(seed)
    p = W.random_small(rng, hx=hx, hy=hy, L=L, grid=16, lossless=lossless)
    gpu = ctx.solve([p])[0]
    ref = O.solve(p)
    dR, dT, dtot = _cmp(gpu, ref)
    assert max(dR, dT, dtot) < TOL, (dR, dT, dtot)
    if lossless:
        assert np.abs(gpu[2].sum(axis=1) - 1.0).max() < 1e-10


def test_solve_zero_layers(ctx):
    rng = np.random.default_rng(9)
    p = W.random_small(rng, hx=1, hy=1, L=1, grid=8)
    p.er = p.er[:0]
    p.ur = p.ur[:0]
    p.thick = p.thick[:0]
    gpu = ctx.solve([p])[0]
    ref = O.solve(p)
    assert max(_cmp(gpu, ref)) < TOL


def test_solve_batch_mixed_shapes(ctx):
    rng = np.random.default_rng(11)
    probs = [W.random_small(rng, hx=1, hy=1, L=2, grid=8),
             W.random_small(rng, hx=2, hy=1, L=1, grid=8),
             W.random_small(rng, hx=1, hy=1, L=2, grid=8)]
    out = ctx.solve(probs)
    for p, g in zip(probs, out):
        assert max(_cmp(g, O.solve(p))) < TOL


def test_c1_vs_oracle(ctx):
    p = W.c1_single()
    gpu = ctx.solve([p])[0]
    ref = O.solve(p)
    dR, dT, dtot = _cmp(gpu, ref)
    assert max(dR, dT, dtot) < TOL, (dR, dT, dtot)


def test_c1_lossless_energy(ctx):
    p = W.c1_single()
    p.er = p.er.real.astype(np.complex128)
    gpu = ctx.solve([p])[0]
    assert abs(gpu[2][0].sum() - 1.0) < 1e-10
    ref = O.solve(p)
    assert max(_cmp(gpu, ref)) < TOL


def test_split_form_error_vs_fp64(ctx):
    """Split-precision form (Ozaki-sliced FP16 GEMMs for the layer S-matrix
    and Redheffer products; FP64 Omega^2, V, eigensolver and inverses) against
    the FP64 form on the same structure; the error is reported, and bounded by
    the 1e-6 target."""
    p = W.c1_single(grid=128, h=3)
    f = ctx.solve([p])[0]
    s = ctx.solve([p], precision=G.RCWA_SPLIT)[0]
    dR = np.abs(f[0] - s[0]).max()
    dT = np.abs(f[1] - s[1]).max()
    print("split vs fp64: max |dR| = %.3e, max |dT| = %.3e" % (dR, dT))
    assert dR < 1e-6 and dT < 1e-6


def test_precision_switches_on_one_context(ctx):
    """FP64, split, FP64, split on one context: the split form's slice
    workspaces (layer stage and fold stream) are allocated on demand and
    released by the next FP64 solve; every repeat gives the same results."""
    rng = np.random.default_rng(53)
    probs = [W.random_small(rng, hx=3, hy=3, L=4, grid=16) for _ in range(3)]
    f1 = ctx.solve(probs)
    s1 = ctx.solve(probs, precision=G.RCWA_SPLIT)
    f2 = ctx.solve(probs)
    s2 = ctx.solve(probs, precision=G.RCWA_SPLIT)
    for a, b in zip(f1, f2):
        assert max(_cmp(a, b)) == 0.0
    for a, b in zip(s1, s2):
        assert max(_cmp(a, b)) == 0.0
    for a, b in zip(f1, s1):
        assert max(_cmp(a, b)) < 1e-6


def test_solve_batch_mixed_layer_counts(ctx):
    """One chunk with different layer counts: the fold runs on a shrinking
    prefix of the jobs (layer-major layout)."""
    rng = np.random.default_rng(31)
    probs = [W.random_small(rng, hx=1, hy=2, L=L, grid=8) for L in (1, 4, 2, 3, 0, 4)]
    probs[4].er = probs[4].er[:0]
    probs[4].ur = probs[4].ur[:0]
    probs[4].thick = probs[4].thick[:0]
    out = ctx.solve(probs)
    for p, g in zip(probs, out):
        assert max(_cmp(g, O.solve(p))) < TOL


def test_line_grating_sweep_vs_oracle(ctx):
    """The benchmark's structure at a testable size: the trapezoid line
    grating of 20 slices over 5 dispersive materials, theta = 65 deg, TE and
    TM, 7x7 harmonics on 64x64 grids, at wavelengths across the 250-800 nm
    sweep (including the substrate cutoff region near 289 nm), solved as one
    batch and compared per order with the oracle."""
    probs = W.c2_sweep(indices=[0, 14, 70, 199], grid=64, h=3)
    out = ctx.solve(probs)
    for p, gpu in zip(probs, out):
        ref = O.solve(p)
        dR, dT, dtot = _cmp(gpu, ref)
        assert max(dR, dT, dtot) < TOL, (p.lam0, dR, dT, dtot)


def _oracle_amplitudes(p):
    """Per-order r, t of the oracle's global S-matrix (the numpy
    restatement's source and field steps, rcwa.cpp:469-561, W_ref = I)."""
    _, _, _, Sg = O.solve(p, return_smatrix=True)
    k, Kx, Ky, Kzr, Kzt = O.wave_vectors(p.Lx, p.Ly, p.lam0, p.theta, p.phi, p.hx, p.hy,
                                         p.er_ref, p.ur_ref, p.er_trn, p.ur_trn)
    N = Kx.size
    I = np.eye(2 * N, dtype=np.complex128)
    out = []
    for a, b in p.pols:
        csrc = NR.source_coeff(p.theta, a, b, k, N, I)
        out.append((NR.fields(Sg, csrc, Kx, Ky, Kzr, I, 0).reshape(3, N),
                    NR.fields(Sg, csrc, Kx, Ky, Kzt, I, 1).reshape(3, N)))
    return out


@pytest.mark.parametrize("seed,lossless", [(41, True), (42, False)])
def test_amplitudes_vs_oracle(ctx, seed, lossless):
    """Per-order complex amplitudes r, t (the fields of rcwa.cpp:526-561)
    within 1e-9 of the oracle's, phases included, through the batched solve
    and through the serving form."""
    rng = np.random.default_rng(seed)
    probs = [W.random_small(rng, hx=2, hy=2, L=3, grid=16, lossless=lossless) for _ in range(2)]
    out = ctx.solve(probs, fields=True)
    with G.Server(0, n_workers=1) as srv:
        tickets = [srv.submit(p) for p in probs]
        served = [srv.wait(t, fields=True) for t in tickets]
    for p, g, s in zip(probs, out, served):
        ref = _oracle_amplitudes(p)
        for j, (r_ref, t_ref) in enumerate(ref):
            assert np.abs(g[3][j] - r_ref).max() < TOL
            assert np.abs(g[4][j] - t_ref).max() < TOL
            assert np.abs(s[4][j] - r_ref).max() < TOL
            assert np.abs(s[5][j] - t_ref).max() < TOL


def test_layer_batches_split_at_two_per_cu(ctx):
    """More layers of n >= 256 than two per CU (13x13 harmonics, n = 338;
    35 structures x 15 layers + one of 2 = 527 > 2 x 256) are split into
    equal layer batches, with the fold of the first running beside the
    second: every result equals the solve of its structure in a chunk small
    enough for one batch, and the two-layer structure equals the oracle."""
    rng = np.random.default_rng(47)
    probs = [W.random_small(rng, hx=6, hy=6, L=2, grid=16)]
    probs += [W.random_small(rng, hx=6, hy=6, L=15, grid=16) for _ in range(35)]
    together = ctx.solve(probs)
    ctx.set_chunk(12)  # 180 layers per chunk: one batch
    try:
        apart = ctx.solve(probs)
    finally:
        ctx.set_chunk(0)
    for g, h in zip(together, apart):
        assert max(_cmp(g, h)) < 1e-12
    assert max(_cmp(together[0], O.solve(probs[0]))) < TOL


def test_c2_full_size_lossless_energy(ctx):
    """The benchmark's structure at full size (21x21 harmonics, n = 882, 20
    layers on 512^2 grids) at two wavelengths of its sweep, every permittivity
    made real: R + T = 1 to 1e-10 for TE and TM, and no negative order.  A
    size-independent check at the size where the CPU oracle would take about a
    quarter of an hour per solve."""
    probs = W.c2_sweep(indices=[0, 137])
    for p in probs:
        p.er = p.er.real.astype(np.complex128)
    out = ctx.solve(probs)
    for R, T, tot in out:
        dev = np.abs(np.asarray(tot).sum(axis=1) - 1.0).max()
        print("R + T - 1: %.3e, min order %.3e" % (dev, min(R.min(), T.min())))
        assert R.min() > -1e-12 and T.min() > -1e-12
        assert dev < 1e-10


def test_c2_full_size_split_and_sweep_plan(ctx):
    """Full size again (21x21 harmonics, 20 layers), two wavelengths: the
    split-precision form stays within the 1e-6 target of the FP64 form, and
    the sweep plan (convolution matrices mixed from material masks) equals the
    direct solve on the rasterised grids to 1e-11."""
    idx = [20, 150]
    probs = W.c2_sweep(indices=idx)
    f = ctx.solve(probs)
    s = ctx.solve(probs, precision=G.RCWA_SPLIT)
    for a, b in zip(f, s):
        assert max(_cmp(a, b)) < 1e-6
    material, thick, lams, eps = W.c2_structure()
    th = math.radians(65.0)
    srcs = [[(lams[i], th, 0.0, 1.0, 0.0), (lams[i], th, 0.0, 0.0, 1.0)] for i in idx]
    sw = ctx.sweep(material, thick, 120.0, 120.0, eps.shape[1], 10, 10)
    try:
        pres = sw.run(srcs, eps[idx])
    finally:
        sw.close()
    for a, b in zip(f, pres):
        assert max(_cmp(a, b)) < 1e-11

"""GPU kernels against numpy / the oracle: ZGEMM on the FP64 matrix cores,
batched Gauss-Jordan inverse, batched eigendecomposition, ConvMat."""
import numpy as np
import pytest

import oracle_py as O
import rcwa_amd as G

pytestmark = pytest.mark.gpu


def rnd(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (16, 16, 4), (17, 33, 5), (64, 64, 64),
                                   (100, 70, 130), (242, 242, 242)])
def test_zgemm_fp64(ctx, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = [rnd(rng, m, k) for _ in range(3)]
    B = [rnd(rng, k, n) for _ in range(3)]
    C = ctx.zgemm(A, B)
    for a, b, c in zip(A, B, C):
        ref = a @ b
        err = np.abs(c - ref).max() / (np.abs(a).max() * np.abs(b).max() * k)
        assert err < 1e-15, err


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 100, 242])
def test_inverse(ctx, n):
    rng = np.random.default_rng(n)
    A = [rnd(rng, n, n) + 2 * np.sqrt(n) * np.eye(n) for _ in range(4)]
    X = ctx.inverse(A)
    for a, x in zip(A, X):
        err = np.abs(a @ x - np.eye(n)).max()
        assert err < 1e-12 * n, err
        assert np.abs(x - np.linalg.inv(a)).max() / np.abs(x).max() < 1e-12


@pytest.mark.parametrize("n", [1100, 2100])
def test_inverse_large(ctx, n):
    """Register panels with two (n <= 2048) and four (n <= 4096) rows per
    thread; n = 1250 is the 25x25-harmonic size of the serving stream."""
    rng = np.random.default_rng(n)
    A = [rnd(rng, n, n) + 2 * np.sqrt(n) * np.eye(n) for _ in range(2)]
    A[1][[0, 5]] = A[1][[5, 0]]   # force a swap at the first pivot
    A[1][0, 0] = 0
    X = ctx.inverse(A)
    for a, x in zip(A, X):
        assert np.abs(a @ x - np.eye(n)).max() < 1e-12 * n
        assert np.abs(x - np.linalg.inv(a)).max() / np.abs(x).max() < 1e-11


@pytest.mark.parametrize("n", [40, 300, 882])
def test_inverse_random_pivots(ctx, n):
    """Unshifted Gaussian matrices: nearly every column pivots, so each panel's
    row swaps (up to 2 x 16 rows, applied to the other columns as one
    permutation) are exercised; errors scale with the condition number."""
    rng = np.random.default_rng(7 + n)
    A = [rnd(rng, n, n) for _ in range(3)]
    X = ctx.inverse(A)
    for a, x in zip(A, X):
        ref = np.linalg.inv(a)
        cond = np.linalg.cond(a)
        assert np.abs(a @ x - np.eye(n)).max() < 1e-14 * n * cond
        assert np.abs(x - ref).max() / np.abs(ref).max() < 1e-14 * n * cond


def test_inverse_needs_pivoting(ctx):
    # zero leading entry forces a row swap at the first step
    A = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 9]], np.complex128) + 0.5j * np.eye(3)
    X = ctx.inverse([A])[0]
    assert np.allclose(X, np.linalg.inv(A), atol=1e-13)


def _match_eigs(w, ref):
    ref = list(ref)
    errs = []
    for v in w:
        i = int(np.argmin([abs(v - r) for r in ref]))
        errs.append(abs(v - ref[i]))
        ref.pop(i)
    return max(errs)


@pytest.mark.parametrize("n", [2, 3, 8, 40, 98, 242])
def test_eig_residual_and_values(ctx, n):
    rng = np.random.default_rng(100 + n)
    A = [rnd(rng, n, n) for _ in range(3)]
    w, V = ctx.eig(A)
    for a, wi, vi in zip(A, w, V):
        nrm = np.abs(a).max()
        res = np.abs(a @ vi - vi * wi[None, :]).max() / nrm
        assert res < 1e-12 * n, res
        assert np.allclose(np.linalg.norm(vi, axis=0), 1.0, atol=1e-12)
        assert _match_eigs(wi, np.linalg.eigvals(a)) < 1e-10 * nrm


@pytest.mark.parametrize("n,cuts", [(60, [20]), (242, [80, 160]), (300, [7, 150, 290])])
def test_eig_block_triangular(ctx, n, cuts):
    """Exact zero blocks below the diagonal survive the Hessenberg reduction,
    so the QR iteration splits in the middle of the matrix: the active block
    [il, iu] moves up and down, and updates deferred for rows above il and
    columns right of iu must still land."""
    rng = np.random.default_rng(n)
    a = rnd(rng, n, n)
    edges = [0] + cuts + [n]
    for lo, hi in zip(edges[:-1], edges[1:]):
        a[hi:, lo:hi] = 0
    w, V = ctx.eig([a, a.T.copy()])
    for m, wi, vi in zip([a, a.T], w, V):
        nrm = np.abs(m).max()
        res = np.abs(m @ vi - vi * wi[None, :]).max() / nrm
        assert res < 1e-12 * n, res
        assert _match_eigs(wi, np.linalg.eigvals(m)) < 1e-10 * nrm


@pytest.mark.parametrize("cap", [1, 3000])
def test_eig_small_rotation_log(cap):
    """A rotation log too small for the whole iteration: the sweep kernel
    applies it itself whenever it fills (cap 1 is raised to one sweep's worth,
    so that happens before nearly every sweep), and what is left is applied
    afterwards.  Same eigenpairs as with the full log."""
    n = 150
    rng = np.random.default_rng(cap)
    a = rnd(rng, n, n)
    a[100:, :100] = 0       # a split in the middle as well
    c = G.Context(0)
    try:
        c.set_qr_log(cap)
        w, V = c.eig([a, a.T.copy()])
    finally:
        c.close()
    for m, wi, vi in zip([a, a.T], w, V):
        nrm = np.abs(m).max()
        res = np.abs(m @ vi - vi * wi[None, :]).max() / nrm
        assert res < 1e-12 * n, res
        assert _match_eigs(wi, np.linalg.eigvals(m)) < 1e-10 * nrm


@pytest.mark.parametrize("mode", ["", "pipe", "single", "oneshot"])
def test_eig_sweep_forms(ctx, monkeypatch, mode):
    """Every form of the QR sweep (the default multishift one with aggressive
    early deflation and eight chained sweeps per pass, the pipelined one with
    two chained sweeps and one shift per pass, the pipelined one with one
    sweep, the one-shot logged sweep used for matrices of 2^31 bytes or more)
    on a matrix that splits in the middle, with a rotation log small enough
    to fill and resume."""
    if mode:
        monkeypatch.setenv("RCWA_QR_SWEEP", mode)
    n = 170
    rng = np.random.default_rng(17)
    a = rnd(rng, n, n)
    a[110:, :110] = 0
    c = G.Context(0)
    try:
        c.set_qr_log(2000)
        w, V = c.eig([a, a.T.copy()])
    finally:
        c.close()
    for m, wi, vi in zip([a, a.T], w, V):
        nrm = np.abs(m).max()
        res = np.abs(m @ vi - vi * wi[None, :]).max() / nrm
        assert res < 1e-12 * n, res
        assert _match_eigs(wi, np.linalg.eigvals(m)) < 1e-10 * nrm


def test_eig_workload_size(ctx):
    """n = 882: the Omega^2 size of 21 x 21 harmonics."""
    n = 882
    rng = np.random.default_rng(882)
    A = [rnd(rng, n, n) for _ in range(2)]
    w, V = ctx.eig(A)
    for a, wi, vi in zip(A, w, V):
        nrm = np.abs(a).max()
        res = np.abs(a @ vi - vi * wi[None, :]).max() / nrm
        assert res < 1e-12 * n, res
        assert _match_eigs(wi, np.linalg.eigvals(a)) < 1e-9 * nrm


def test_eig_large_sampled_residual(ctx):
    """n = 2100 (the one-workgroup sweep kernel and the log apply at large n,
    register Gauss-Jordan with two rows per thread in the caller): residual
    on sampled eigenpairs, eigenvalues against LAPACK on a smaller copy
    would be too slow here, so their sum is checked against the trace."""
    n = 2100
    rng = np.random.default_rng(2100)
    a = rnd(rng, n, n)
    w, V = ctx.eig([a])
    w, V = w[0], V[0]
    cols = np.arange(0, n, 97)
    res = np.abs(a @ V[:, cols] - V[:, cols] * w[None, cols]).max() / np.abs(a).max()
    assert res < 1e-12 * n, res
    assert abs(w.sum() - np.trace(a)) < 1e-9 * np.abs(a).max() * n


def test_eig_degenerate_diagonal(ctx):
    # a uniform layer gives a diagonal Omega^2 with every eigenvalue twice
    d = np.repeat(np.arange(1, 6) - 3.5 + 0.1j, 2)
    w, V = ctx.eig([np.diag(d)])
    assert _match_eigs(w[0], d) < 1e-14
    assert np.linalg.cond(V[0]) < 1e3


@pytest.mark.parametrize("grid,hx,hy", [((8, 8), 1, 1), ((8, 6), 2, 1), ((16, 12), 2, 3),
                                        ((64, 64), 5, 5)])
def test_convmat_vs_oracle(ctx, grid, hx, hy):
    rng = np.random.default_rng(7)
    Ny, Nx = grid
    g = rnd(rng, 2, Ny, Nx)
    for layer in range(2):
        c = ctx.convmat(g, layer, hx, hy)
        ref = O.convmat(g[layer], hx, hy)
        assert np.abs(c - ref).max() < 1e-14 * np.abs(g).max()


@pytest.mark.parametrize("m,n,k", [(16, 16, 16), (70, 33, 100), (242, 242, 242),
                                   (16, 40, 40000)])
def test_zgemm_split_accuracy(ctx, m, n, k):
    """Ozaki scheme with six 7-bit fixed-point int8 slices, the slice products
    of weight 2^-49 and above accumulated exactly in int32 by weight group:
    error below about 2^-37 of max|a_row| max|b_col| per k term (the dropped
    slice products).  k = 40000 crosses the 2^15-k chunks the int32 sums are
    cut into."""
    rng = np.random.default_rng(m + n + k)
    A = [rnd(rng, m, k) * np.exp(rng.uniform(-3, 3, (m, 1))) for _ in range(2)]
    B = [rnd(rng, k, n) for _ in range(2)]
    C = ctx.zgemm(A, B, split=True)
    for a, b, c in zip(A, B, C):
        ref = a @ b
        bound = np.abs(a).max(axis=1, keepdims=True) * np.abs(b).max(axis=0, keepdims=True) * k
        err = (np.abs(c - ref) / bound).max()
        assert err < 5e-11, err

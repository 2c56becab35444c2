"""Pins the CPU FP64 oracle (oracle/rcwa_oracle.cpp).

The reference (rcwa3d) ships no tests, golden vectors or fixtures, and it
cannot be built here (Eigen3 is absent), so every oracle function is pinned
by what the mathematics fixes, chosen so that a dropped term, a wrong sign or
index, or a transposed operand fails at least one check:

  ConvMat            brute-force DFT sum; uniform grid -> c I; Toeplitz
                     structure; zero order = grid mean; linearity
  ComputeWaveVectors explicit scalar formulas and branch signs
  GapMedium          closed form, W0 = I, V0 2x2-block diagonal
  eigensolver        residual, eigenvalues vs LAPACK zgeev, degenerate case
  SMatrixLayer       t = 0 -> (0, I); gap-medium layer -> (0, X);
                     splitting t = t1 + t2 gives the same S
  RedhefferProduct   identity element (exact), associativity, 1x1 formula
  end to end         Fresnel at one interface (TE, TM, oblique); uniform
                     multilayers vs a transfer-matrix formula; R + T = 1 for
                     lossless stacks; mirror symmetry; agreement with the
                     independent numpy/LAPACK restatement (oracle/np_rcwa.py)
"""
import numpy as np
import pytest

import np_rcwa as NR
import oracle_py as O
import workloads as W


def rnd(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


# ------------------------------------------------------------------ ConvMat
def brute_convmat(grid, hx, hy):
    Ny, Nx = grid.shape
    P, Q = 2 * hx + 1, 2 * hy + 1
    C = np.zeros((P * Q, P * Q), complex)
    ys, xs = np.mgrid[0:Ny, 0:Nx]
    for m in range(P * Q):
        for n in range(P * Q):
            dp = (m // Q - hx) - (n // Q - hx)
            dq = (m % Q - hy) - (n % Q - hy)
            if not (0 <= dp + Nx // 2 < Nx and 0 <= dq + Ny // 2 < Ny):
                continue
            ph = np.exp(-2j * np.pi * (dp * xs / Nx + dq * ys / Ny))
            C[m, n] = (grid * ph).sum() / (Nx * Ny)
    return C


@pytest.mark.parametrize("shape,hx,hy", [((4, 4), 1, 1), ((6, 8), 2, 1), ((5, 7), 1, 2),
                                         ((8, 8), 2, 2), ((3, 3), 2, 2)])
def test_convmat_brute_force(shape, hx, hy):
    rng = np.random.default_rng(sum(shape) + hx)
    g = rnd(rng, *shape)
    C = O.convmat(g, hx, hy)
    assert np.abs(C - brute_convmat(g, hx, hy)).max() < 1e-13


def test_convmat_structure():
    rng = np.random.default_rng(1)
    g = rnd(rng, 16, 16)
    C = O.convmat(g, 2, 2)
    assert np.allclose(O.convmat(np.full((16, 16), 3 - 2j), 2, 2), (3 - 2j) * np.eye(25))
    assert abs(C[12, 12] - g.mean()) < 1e-14                     # zero order
    # Toeplitz: depends only on (dp, dq)
    assert abs(C[0, 1] - C[5, 6]) < 1e-14 and abs(C[6, 0] - C[7, 1]) < 1e-14
    g2 = rnd(rng, 16, 16)
    assert np.allclose(O.convmat(2 * g + g2, 2, 2), 2 * C + O.convmat(g2, 2, 2), atol=1e-13)
    # real symmetric grid -> Hermitian convolution matrix
    r = rng.random((16, 16))
    assert np.allclose(O.convmat(r, 2, 2), O.convmat(r, 2, 2).conj().T, atol=1e-14)


# ------------------------------------------------------------------ wave vectors
def test_wave_vectors_formulas():
    Lx, Ly, lam, th, ph = 400.0, 300.0, 633.0, 0.4, 0.3
    er, ur, et, ut = 1.5 - 0.1j, 1.2, 2.25, 1.0
    k, Kx, Ky, Kzr, Kzt = O.wave_vectors(Lx, Ly, lam, th, ph, 2, 1, er, ur, et, ut)
    n = np.sqrt(er * ur)
    assert np.allclose(k, [n * np.sin(th) * np.cos(ph), n * np.sin(th) * np.sin(ph), n * np.cos(th)])
    m = 7                      # ix = 2 (p = 0), iy = 1 (q = 0) with Q = 3
    assert abs(Kx[m] - k[0]) < 1e-15 and abs(Ky[m] - k[1]) < 1e-15
    assert abs(Kx[0] - (k[0] + 2 * lam / Lx)) < 1e-14           # index 0 has the largest kx
    assert abs(Ky[0] - (k[1] + lam / Ly)) < 1e-14
    for i in range(15):
        zr = -np.conj(NR.usqrt(np.conj(ur) * np.conj(er) - Kx[i] ** 2 - Ky[i] ** 2))
        assert abs(Kzr[i] - zr) < 1e-14
    # lossless branches: propagating Re Kz_ref < 0, evanescent Im Kz_ref > 0
    k, Kx, Ky, Kzr, Kzt = O.wave_vectors(Lx, Ly, lam, th, 0.0, 2, 1, 1.0, 1.0, 2.25, 1.0)
    prop = np.abs(Kx) ** 2 + np.abs(Ky) ** 2 < 1
    assert prop.any() and (~prop).any()
    assert (Kzr[prop].real < 0).all() and (Kzr[~prop].imag > 0).all()
    assert (Kzt.real >= 0).all()


def test_wave_vectors_match_numpy():
    args = (500.0, 450.0, 700.0, 0.2, 0.7, 3, 2, 1.0 + 0j, 1.0, 2.1 - 0.3j, 1.0)
    a = O.wave_vectors(*args)
    b = NR.wave_vectors(*args)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() < 1e-14


# ------------------------------------------------------------------ eigensolver
@pytest.mark.parametrize("n", [1, 2, 5, 30, 100])
def test_eig_residual_and_lapack(n):
    rng = np.random.default_rng(n)
    A = rnd(rng, n, n)
    w, V, it = O.eig(A)
    assert np.abs(A @ V - V * w).max() < 1e-12 * max(1, n) * np.abs(A).max()
    assert np.allclose(np.linalg.norm(V, axis=0), 1)
    ref = np.sort_complex(np.linalg.eigvals(A))
    assert np.abs(np.sort_complex(w) - ref).max() < 1e-11 * np.abs(A).max()


def test_eig_degenerate_and_triangular():
    d = np.repeat(np.array([1.0, -2.0 + 0.5j, 3.0j]), 2)
    w, V, _ = O.eig(np.diag(d))
    assert np.allclose(np.sort_complex(w), np.sort_complex(d))
    assert np.linalg.cond(V) < 10
    # a similarity transform of a matrix with double eigenvalues keeps them
    rng = np.random.default_rng(3)
    S = rnd(rng, 6, 6) + 4 * np.eye(6)
    A = S @ np.diag(d) @ np.linalg.inv(S)
    w, V, _ = O.eig(A)
    assert np.abs(np.sort_complex(w) - np.sort_complex(d)).max() < 1e-10
    assert np.abs(A @ V - V * w).max() < 1e-11 * np.abs(A).max()


def test_hessenberg_is_similarity():
    rng = np.random.default_rng(4)
    A = rnd(rng, 12, 12)
    H, Q = O.hessenberg(A)
    assert np.abs(np.tril(H, -2)).max() == 0
    assert np.allclose(Q.conj().T @ Q, np.eye(12), atol=1e-14)
    assert np.allclose(Q @ H @ Q.conj().T, A, atol=1e-13)


# ------------------------------------------------------------------ layer S-matrix
def _layer_inputs(rng, hx=1, hy=1, grid=8):
    p = W.random_small(rng, hx=hx, hy=hy, L=1, grid=grid)
    k, Kx, Ky, Kzr, Kzt = O.wave_vectors(p.Lx, p.Ly, p.lam0, p.theta, p.phi, hx, hy, 1, 1, 2.25, 1)
    return p, Kx, Ky, O.convmat(p.er[0], hx, hy), O.convmat(p.ur[0], hx, hy), 2 * np.pi / p.lam0


def test_layer_zero_thickness():
    rng = np.random.default_rng(5)
    p, Kx, Ky, erc, urc, k0 = _layer_inputs(rng)
    S11, S12 = O.smatrix_layer(erc, urc, Kx, Ky, k0, 0.0)
    n = S11.shape[0]
    assert np.abs(S11).max() < 1e-12 and np.abs(S12 - np.eye(n)).max() < 1e-12


def test_layer_equal_to_gap_medium():
    rng = np.random.default_rng(6)
    p, Kx, Ky, _, _, k0 = _layer_inputs(rng)
    N = len(Kx)
    t = 123.0
    S11, S12 = O.smatrix_layer(np.eye(N), np.eye(N), Kx, Ky, k0, t)
    lam0 = 1j * np.conj(NR.usqrt(1 - Kx ** 2 - Ky ** 2))
    X = np.exp(-k0 * t * np.concatenate([lam0, lam0]))
    assert np.abs(S11).max() < 1e-12
    assert np.abs(S12 - np.diag(X)).max() < 1e-12


def test_layer_thickness_split():
    rng = np.random.default_rng(7)
    p, Kx, Ky, erc, urc, k0 = _layer_inputs(rng, 2, 1)
    S = O.smatrix_layer(erc, urc, Kx, Ky, k0, 70.0)
    S1 = O.smatrix_layer(erc, urc, Kx, Ky, k0, 30.0)
    S2 = O.smatrix_layer(erc, urc, Kx, Ky, k0, 40.0)
    comb = O.redheffer((S1[0], S1[1], S1[1], S1[0]), (S2[0], S2[1], S2[1], S2[0]))
    assert np.abs(comb[0] - S[0]).max() < 1e-10
    assert np.abs(comb[1] - S[1]).max() < 1e-10
    assert np.abs(comb[2] - S[1]).max() < 1e-10


def test_layer_matches_numpy():
    rng = np.random.default_rng(8)
    p, Kx, Ky, erc, urc, k0 = _layer_inputs(rng, 2, 2, 12)
    S11, S12 = O.smatrix_layer(erc, urc, Kx, Ky, k0, 150.0)
    W0, V0 = NR.gap_medium(Kx, Ky)
    R = NR.smatrix_layer(erc, urc, Kx, Ky, k0, 150.0, W0, V0)
    assert np.abs(S11 - R[0]).max() < 1e-11 and np.abs(S12 - R[1]).max() < 1e-11


# ------------------------------------------------------------------ Redheffer
def test_redheffer_identity_and_assoc():
    rng = np.random.default_rng(9)
    n = 6
    mk = lambda: tuple(0.4 * rnd(rng, n, n) for _ in range(4))
    A, B, C = mk(), mk(), mk()
    Id = NR.smatrix_init(n)
    left = O.redheffer(Id, A)
    assert all(np.array_equal(x, y) for x, y in zip(left, A))   # exact
    AB_C = O.redheffer(O.redheffer(A, B), C)
    A_BC = O.redheffer(A, O.redheffer(B, C))
    assert all(np.abs(x - y).max() < 1e-12 for x, y in zip(AB_C, A_BC))


def test_redheffer_scalar():
    a = [np.array([[v]]) for v in (0.3 + 0.1j, 0.8, 0.7 - 0.2j, -0.25j)]
    b = [np.array([[v]]) for v in (0.1 - 0.4j, 0.6j, 0.5, 0.2)]
    S = O.redheffer(a, b)
    A11, A12, A21, A22 = (x[0, 0] for x in a)
    B11, B12, B21, B22 = (x[0, 0] for x in b)
    d = 1 - B11 * A22
    assert abs(S[0][0, 0] - (A11 + A12 * B11 * A21 / d)) < 1e-15
    assert abs(S[1][0, 0] - A12 * B12 / d) < 1e-15
    assert abs(S[2][0, 0] - B21 * A21 / d) < 1e-15
    assert abs(S[3][0, 0] - (B22 + B21 * A22 * B12 / d)) < 1e-15


# ------------------------------------------------------------------ end to end
def fresnel(n1, n2, th, pol):
    c1 = np.cos(th)
    c2 = np.sqrt(1 - (n1 * np.sin(th) / n2) ** 2 + 0j)
    if pol == "TE":
        r = (n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)
    else:
        r = (n2 * c1 - n1 * c2) / (n2 * c1 + n1 * c2)
    return abs(r) ** 2


@pytest.mark.parametrize("th", [0.0, 0.35, 1.0])
def test_single_interface_fresnel(th):
    er_t = 2.25
    g = np.ones((0, 8, 8), complex)
    p = W.Problem(g, g, [], 200.0, 200.0, 600.0, th, 0.0, [(1, 0), (0, 1)], 1, 1,
                  1.0, 1.0, er_t, 1.0)
    R, T, tot = O.solve(p)
    for i, pol in enumerate(("TE", "TM")):
        Rf = fresnel(1.0, 1.5, th, pol)
        assert abs(tot[i][0] - Rf) < 1e-12
        assert abs(tot[i][1] - (1 - Rf)) < 1e-12


def tmm(ns, ds, n_in, n_out, th, lam, pol):
    """Characteristic-matrix reflectance/transmittance of a lossless stack."""
    s = n_in * np.sin(th)
    cos = lambda n: np.sqrt(1 - (s / n) ** 2 + 0j)
    eta = (lambda n: n * cos(n)) if pol == "TE" else (lambda n: n / cos(n))
    M = np.eye(2, dtype=complex)
    for n, d in zip(ns, ds):
        dl = 2 * np.pi / lam * n * cos(n) * d
        e = eta(n)
        M = M @ np.array([[np.cos(dl), 1j * np.sin(dl) / e], [1j * e * np.sin(dl), np.cos(dl)]])
    e0, es = eta(n_in), eta(n_out)
    B, C = M @ np.array([1, es])
    r = (e0 * B - C) / (e0 * B + C)
    t = 2 * e0 / (e0 * B + C)
    return abs(r) ** 2, (es.real / e0.real) * abs(t) ** 2


@pytest.mark.parametrize("th", [0.0, 0.5])
def test_uniform_multilayer_vs_tmm(th):
    eps = [4.0, 2.1, 6.5]
    ds = [120.0, 80.0, 55.0]
    g = np.stack([np.full((8, 8), e, complex) for e in eps])
    p = W.Problem(g, np.ones_like(g), ds, 250.0, 250.0, 633.0, th, 0.0, [(1, 0), (0, 1)], 1, 1,
                  1.0, 1.0, 2.25, 1.0)
    R, T, tot = O.solve(p)
    for i, pol in enumerate(("TE", "TM")):
        Rt, Tt = tmm(np.sqrt(eps), ds, 1.0, 1.5, th, 633.0, pol)
        assert abs(tot[i][0] - Rt) < 1e-11, (pol, tot[i][0], Rt)
        assert abs(tot[i][1] - Tt) < 1e-11


@pytest.mark.parametrize("seed", range(4))
def test_lossless_energy_conservation(seed):
    rng = np.random.default_rng(100 + seed)
    p = W.random_small(rng, hx=1, hy=1, L=2, grid=8, lossless=True)
    R, T, tot = O.solve(p)
    assert np.abs(tot.sum(axis=1) - 1).max() < 1e-10


def test_mirror_symmetry_normal_incidence():
    p = W.c1_single(grid=64, h=2)
    p.er = p.er.real.astype(complex)
    R, T, tot = O.solve(p)
    assert np.abs(R[0] - R[0][:, ::-1]).max() < 1e-12
    assert np.abs(R[0] - R[0][::-1, :]).max() < 1e-12
    assert abs(tot[0].sum() - 1) < 1e-10


@pytest.mark.parametrize("seed,hx,hy,L,lossless", [(0, 1, 1, 2, False), (1, 2, 1, 3, True),
                                                   (2, 0, 2, 1, False), (3, 2, 2, 2, False)])
def test_oracle_vs_numpy_restatement(seed, hx, hy, L, lossless):
    rng = np.random.default_rng(seed)
    p = W.random_small(rng, hx=hx, hy=hy, L=L, grid=10, lossless=lossless,
                       uniform_mu=(seed % 2 == 0))
    R, T, tot = O.solve(p)
    for pol in range(len(p.pols)):
        Rn, Tn, Rt, Tt = NR.solve(*p.np_args(pol))
        assert np.abs(R[pol] - Rn).max() < 1e-12
        assert np.abs(T[pol] - Tn).max() < 1e-12

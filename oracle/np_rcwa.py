"""Independent numpy/LAPACK restatement of rcwa3d's FP64 solve path.

TEST INFRASTRUCTURE ONLY.  This module is a checker: it is imported by the
test suite (and by the fixture generator under tests/golden/) to cross-check
the C++ oracle in oracle/rcwa_oracle.cpp.  No product code path may import it.

It follows the reference function by function but deliberately uses
different numerical building blocks than the C++ oracle (numpy.fft for the
2-D transform, LAPACK zgeev for the eigendecomposition, LAPACK zgesv for
every solve), so that agreement between the two pins both against a
plausible slip in either.

Reference: rcwa3d/src/rcwa.cpp, rcwa3d/include/rcwa.hpp, rcwa3d/src/main.cpp.
"""
import numpy as np


def usqrt(z):
    """rcwa.hpp:24-28 / 37-44: imag == +-0 is forced to +0 before the
    principal square root, so sqrt(-x) = +i*sqrt(x)."""
    z = np.asarray(z, dtype=np.complex128)
    z = np.where(z.imag == 0.0, z.real + 0j, z)
    return np.sqrt(z)


def convmat(grid, hx, hy):
    """rcwa.cpp:6-82.  grid is [Ny][Nx] (row-major layer slice, rcwa.hpp:56).
    Returns the PQ x PQ Toeplitz matrix with index m = ix*Q + iy."""
    grid = np.asarray(grid, dtype=np.complex128)
    Ny, Nx = grid.shape
    F = np.fft.fft2(grid) / (Nx * Ny)                 # rcwa.cpp:24-41
    Fs = np.fft.fftshift(F)                           # rcwa.cpp:44-53 (zero at N//2)
    P, Q = 2 * hx + 1, 2 * hy + 1
    m = np.arange(P * Q)
    px = m // Q - hx                                  # rcwa.cpp:63 (x slow)
    py = m % Q - hy                                   # rcwa.cpp:64 (y fast)
    kx = px[:, None] - px[None, :] + Nx // 2          # rcwa.cpp:71
    ky = py[:, None] - py[None, :] + Ny // 2          # rcwa.cpp:72
    ok = (kx >= 0) & (kx < Nx) & (ky >= 0) & (ky < Ny)
    C = np.zeros((P * Q, P * Q), dtype=np.complex128)
    C[ok] = Fs[ky[ok], kx[ok]]
    return C


def wave_vectors(Lx, Ly, lam0, theta, phi, hx, hy, er_ref, ur_ref, er_trn, ur_trn):
    """rcwa.cpp:85-160.  Returns k_inc(3), Kx, Ky, Kz_ref, Kz_trn (length PQ)."""
    n_inc = np.sqrt(np.complex128(er_ref * ur_ref))   # plain sqrt, rcwa.cpp:91
    k_inc = np.array([n_inc * np.sin(theta) * np.cos(phi),
                      n_inc * np.sin(theta) * np.sin(phi),
                      n_inc * np.cos(theta)], dtype=np.complex128)
    P, Q = 2 * hx + 1, 2 * hy + 1
    kx = k_inc[0] - np.arange(-hx, hx + 1) * (lam0 / Lx)   # rcwa.cpp:106-112
    ky = k_inc[1] - np.arange(-hy, hy + 1) * (lam0 / Ly)   # rcwa.cpp:113-115
    KX = np.tile(kx[None, :], (Q, 1))                      # MeshGrid, rcwa.hpp:218-224
    KY = np.tile(ky[:, None], (1, P))
    Kx = KX.reshape(-1, order="F")                         # column-major flatten
    Ky = KY.reshape(-1, order="F")
    Kz_ref = -np.conj(usqrt(np.conj(ur_ref) * np.conj(er_ref) - Kx ** 2 - Ky ** 2))
    Kz_trn = np.conj(usqrt(np.conj(ur_trn) * np.conj(er_trn) - Kx ** 2 - Ky ** 2))
    return k_inc, Kx, Ky, Kz_ref, Kz_trn


def _blk(a, b, c, d):
    return np.block([[a, b], [c, d]])


def gap_medium(Kx, Ky):
    """rcwa.cpp:163-234."""
    Kz = np.conj(usqrt(1.0 - Kx ** 2 - Ky ** 2))
    Q = _blk(np.diag(Kx * Ky), np.diag(1.0 - Kx ** 2),
             np.diag(Ky ** 2 - 1.0), np.diag(-Kx * Ky))
    lam = np.concatenate([1j * Kz, 1j * Kz])
    W0 = np.eye(2 * Kx.size, dtype=np.complex128)
    V0 = Q / lam[None, :]
    return W0, V0


def smatrix_init(n):
    """rcwa.cpp:237-250."""
    Z = np.zeros((n, n), np.complex128)
    I = np.eye(n, dtype=np.complex128)
    return (Z, I.copy(), I.copy(), Z.copy())


def smatrix_layer(erc, urc, Kx, Ky, k0, t, W0, V0):
    """rcwa.cpp:254-352."""
    KX, KY = np.diag(Kx), np.diag(Ky)
    e_x = np.linalg.solve(erc, KX)
    e_y = np.linalg.solve(erc, KY)
    Pm = _blk(KX @ e_y, urc - KX @ e_x, KY @ e_y - urc, -KY @ e_x)
    u_x = np.linalg.solve(urc, KX)
    u_y = np.linalg.solve(urc, KY)
    Qm = _blk(KX @ u_y, erc - KX @ u_x, KY @ u_y - erc, -KY @ u_x)
    om2 = Pm @ Qm
    lam2, W = np.linalg.eig(om2)                      # LAPACK zgeev
    lam = usqrt(lam2)
    V = (Qm @ W) / lam[None, :]
    WiW0 = np.linalg.solve(W, W0)
    ViV0 = np.linalg.solve(V, V0)
    A = WiW0 + ViV0
    B = WiW0 - ViV0
    X = np.diag(np.exp(-k0 * t * lam))
    AiX = np.linalg.solve(A, X)
    D = A - X @ B @ AiX @ B
    S11 = np.linalg.solve(D, X @ B @ AiX @ A - B)
    S12 = np.linalg.solve(D, X) @ (A - B @ np.linalg.solve(A, B))
    return (S11, S12, S12.copy(), S11.copy())


def _half_space_Q(Kx, Ky, er, ur):
    return _blk(np.diag(Kx * Ky), np.diag(ur * er - Kx ** 2),
                np.diag(Ky ** 2 - ur * er), np.diag(-Kx * Ky)) / ur


def smatrix_reflection(Kx, Ky, Kz_ref, er_ref, ur_ref, W0, V0):
    """rcwa.cpp:354-411."""
    Qr = _half_space_Q(Kx, Ky, er_ref, ur_ref)
    lam = np.concatenate([-1j * Kz_ref, -1j * Kz_ref])
    Vr = Qr / lam[None, :]
    Wr = np.eye(Qr.shape[0], dtype=np.complex128)
    A = np.linalg.solve(W0, Wr) + np.linalg.solve(V0, Vr)
    B = np.linalg.solve(W0, Wr) - np.linalg.solve(V0, Vr)
    AiB = np.linalg.solve(A, B)
    S11 = -AiB
    S12 = np.linalg.solve(A, 2 * np.eye(A.shape[0]))
    S21 = 0.5 * (A - B @ AiB)
    S22 = np.linalg.solve(A.T, B.T).T
    return (S11, S12, S21, S22), Wr


def smatrix_transmission(Kx, Ky, Kz_trn, er_trn, ur_trn, W0, V0):
    """rcwa.cpp:414-466."""
    Qt = _half_space_Q(Kx, Ky, er_trn, ur_trn)
    lam = np.concatenate([1j * Kz_trn, 1j * Kz_trn])
    Vt = Qt / lam[None, :]
    Wt = np.eye(Qt.shape[0], dtype=np.complex128)
    A = np.linalg.solve(W0, Wt) + np.linalg.solve(V0, Vt)
    B = np.linalg.solve(W0, Wt) - np.linalg.solve(V0, Vt)
    AiB = np.linalg.solve(A, B)
    S11 = np.linalg.solve(A.T, B.T).T
    S12 = 0.5 * (A - B @ AiB)
    S21 = np.linalg.solve(A, 2 * np.eye(A.shape[0]))
    S22 = -AiB
    return (S11, S12, S21, S22), Wt


def redheffer(A, B):
    """rcwa.cpp:588-615."""
    A11, A12, A21, A22 = A
    B11, B12, B21, B22 = B
    n = A11.shape[0]
    I = np.eye(n, dtype=np.complex128)
    M1 = I - B11 @ A22
    M2 = I - A22 @ B11
    S11 = A11 + A12 @ np.linalg.solve(M1, B11) @ A21
    S12 = A12 @ np.linalg.solve(M1, B12)
    S21 = B21 @ np.linalg.solve(M2, A21)
    S22 = B22 + B21 @ np.linalg.solve(M2, A22) @ B12
    return (S11, S12, S21, S22)


def source_coeff(theta, pte, ptm, k_inc, PQ, W_ref):
    """rcwa.cpp:469-523."""
    if abs(theta) < 1e-6:
        aTE = np.array([0, 1, 0], dtype=np.complex128)
    else:
        aTE = np.array([-k_inc[1], k_inc[0], 0], dtype=np.complex128)
    aTE = aTE / np.linalg.norm(aTE)
    aTM = np.array([k_inc[1] * aTE[2] - k_inc[2] * aTE[1],
                    k_inc[2] * aTE[0] - k_inc[0] * aTE[2],
                    k_inc[0] * aTE[1] - k_inc[1] * aTE[0]])
    aTM = aTM / np.linalg.norm(aTM)
    pol = pte * aTE + ptm * aTM
    pol = pol / np.linalg.norm(pol)
    esrc = np.zeros(2 * PQ, dtype=np.complex128)
    esrc[PQ // 2] = pol[0]
    esrc[PQ // 2 + PQ] = pol[1]
    return np.linalg.solve(W_ref, esrc)


def fields(S, csrc, Kx, Ky, Kz, W, which):
    """rcwa.cpp:526-561; which = 0 (S11, reflected) or 1 (S21, transmitted)."""
    PQ = Kx.size
    blk = S[0] if which == 0 else S[2]
    e = W @ (blk @ csrc)
    ez = -(Kx * e[:PQ] + Ky * e[PQ:]) / Kz
    return np.concatenate([e, ez])


def efficiencies(hx, hy, r, t, k_inc, Kz_ref, Kz_trn, ur_ref, ur_trn):
    """rcwa.cpp:564-585.  R, T are Q x P (rows = y order)."""
    P, Q = 2 * hx + 1, 2 * hy + 1
    PQ = P * Q
    rr = np.abs(r[:PQ]) ** 2 + np.abs(r[PQ:2 * PQ]) ** 2 + np.abs(r[2 * PQ:]) ** 2
    tt = np.abs(t[:PQ]) ** 2 + np.abs(t[PQ:2 * PQ]) ** 2 + np.abs(t[2 * PQ:]) ** 2
    den = (k_inc[2] / ur_ref).real
    R = ((-Kz_ref / ur_ref).real / den * rr).reshape(Q, P, order="F")
    T = ((Kz_trn / ur_trn).real / den * tt).reshape(Q, P, order="F")
    return R, T, R.sum(), T.sum()


def solve(er, ur, thick, Lx, Ly, lam0, theta, phi, pte, ptm, hx, hy,
          er_ref, ur_ref, er_trn, ur_trn, return_smatrix=False):
    """main.cpp:126-191 call order.  er, ur: [L][Ny][Nx]."""
    er = np.asarray(er, np.complex128)
    ur = np.asarray(ur, np.complex128)
    L = er.shape[0]
    erc = [convmat(er[l], hx, hy) for l in range(L)]
    urc = [convmat(ur[l], hx, hy) for l in range(L)]
    k_inc, Kx, Ky, Kz_ref, Kz_trn = wave_vectors(Lx, Ly, lam0, theta, phi, hx, hy,
                                                 er_ref, ur_ref, er_trn, ur_trn)
    W0, V0 = gap_medium(Kx, Ky)
    n = 2 * Kx.size
    k0 = 2 * np.pi / lam0
    Sd = smatrix_init(n)
    for l in range(L):
        Sd = redheffer(Sd, smatrix_layer(erc[l], urc[l], Kx, Ky, k0, thick[l], W0, V0))
    Sr, Wr = smatrix_reflection(Kx, Ky, Kz_ref, er_ref, ur_ref, W0, V0)
    St, Wt = smatrix_transmission(Kx, Ky, Kz_trn, er_trn, ur_trn, W0, V0)
    Sg = redheffer(redheffer(Sr, Sd), St)
    csrc = source_coeff(theta, pte, ptm, k_inc, Kx.size, Wr)
    r = fields(Sg, csrc, Kx, Ky, Kz_ref, Wr, 0)
    t = fields(Sg, csrc, Kx, Ky, Kz_trn, Wt, 1)
    out = efficiencies(hx, hy, r, t, k_inc, Kz_ref, Kz_trn, ur_ref, ur_trn)
    if return_smatrix:
        return out, Sg
    return out

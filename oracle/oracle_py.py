"""ctypes front end of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker, never as the product.
"""
import ctypes
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = ctypes.CDLL(path)
        _LIB.oracle_last_error.restype = ctypes.c_char_p
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a, shape=None):
    a = np.asarray(a, dtype=np.complex128)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle error %d: %s" % (rc, lib().oracle_last_error().decode()))


def convmat(grid, hx, hy):
    g = np.ascontiguousarray(grid, dtype=np.complex128)
    Ny, Nx = g.shape
    PQ = (2 * hx + 1) * (2 * hy + 1)
    out = np.zeros((PQ, PQ), np.complex128, order="F")
    _check(lib().oracle_convmat(_p(g), Nx, Ny, hx, hy, _p(out)))
    return out


def eig(A):
    A = np.asfortranarray(A, dtype=np.complex128)
    n = A.shape[0]
    w = np.zeros(n, np.complex128)
    V = np.zeros((n, n), np.complex128, order="F")
    it = ctypes.c_int(0)
    _check(lib().oracle_eig(n, _p(A), _p(w), _p(V), ctypes.byref(it)))
    return w, V, it.value


def hessenberg(A):
    H = np.array(A, dtype=np.complex128, order="F")
    n = H.shape[0]
    Q = np.zeros((n, n), np.complex128, order="F")
    _check(lib().oracle_hessenberg(n, _p(H), _p(Q)))
    return H, Q


def wave_vectors(Lx, Ly, lam0, theta, phi, hx, hy, er_ref, ur_ref, er_trn, ur_trn):
    PQ = (2 * hx + 1) * (2 * hy + 1)
    media = np.array([er_ref, ur_ref, er_trn, ur_trn], np.complex128)
    k = np.zeros(3, np.complex128)
    out = [np.zeros(PQ, np.complex128) for _ in range(4)]
    lib().oracle_wave_vectors(ctypes.c_double(Lx), ctypes.c_double(Ly), ctypes.c_double(lam0),
                              ctypes.c_double(theta), ctypes.c_double(phi), hx, hy, _p(media),
                              _p(k), *[_p(o) for o in out])
    return (k, *out)


def smatrix_layer(erc, urc, Kx, Ky, k0, t):
    N = erc.shape[0]
    erc = np.asfortranarray(erc, np.complex128)
    urc = np.asfortranarray(urc, np.complex128)
    Kx = np.ascontiguousarray(Kx, np.complex128)
    Ky = np.ascontiguousarray(Ky, np.complex128)
    S11 = np.zeros((2 * N, 2 * N), np.complex128, order="F")
    S12 = np.zeros((2 * N, 2 * N), np.complex128, order="F")
    _check(lib().oracle_smatrix_layer(N, _p(erc), _p(urc), _p(Kx), _p(Ky), ctypes.c_double(k0),
                                      ctypes.c_double(t), _p(S11), _p(S12)))
    return S11, S12


def redheffer(A, B):
    n = A[0].shape[0]
    A = [np.asfortranarray(x, np.complex128) for x in A]
    B = [np.asfortranarray(x, np.complex128) for x in B]
    S = [np.zeros((n, n), np.complex128, order="F") for _ in range(4)]
    arr = ctypes.c_void_p * 4
    _check(lib().oracle_redheffer(n, arr(*[x.ctypes.data for x in A]),
                                  arr(*[x.ctypes.data for x in B]),
                                  arr(*[x.ctypes.data for x in S])))
    return tuple(S)


def solve(prob, return_smatrix=False, layer_limit=-1):
    """Full solve of a workloads.Problem for every polarisation it carries.
    Returns R, T as [npol][Q][P] arrays and tot as [npol][2] (R_tot, T_tot)."""
    Ly_, Lx_ = prob.er.shape[1], prob.er.shape[2]
    P, Q = 2 * prob.hx + 1, 2 * prob.hy + 1
    npol = len(prob.pols)
    R = np.zeros((npol, P * Q))
    T = np.zeros((npol, P * Q))
    tot = np.zeros((npol, 2))
    pols = np.array(prob.pols, np.float64).reshape(-1)
    media = np.array([prob.er_ref, prob.ur_ref, prob.er_trn, prob.ur_trn], np.complex128)
    n = 2 * P * Q
    Sg = np.zeros((4, n, n), np.complex128) if return_smatrix else None
    d = ctypes.c_double
    _check(lib().oracle_solve(Lx_, Ly_, prob.L, d(prob.Lx), d(prob.Ly), _p(prob.er), _p(prob.ur),
                              _p(prob.thick), d(prob.lam0), d(prob.theta), d(prob.phi), prob.hx,
                              prob.hy, _p(media), npol, _p(pols), _p(R), _p(T), _p(tot),
                              _p(Sg) if Sg is not None else None, layer_limit))
    R = R.reshape(npol, P, Q).transpose(0, 2, 1)
    T = T.reshape(npol, P, Q).transpose(0, 2, 1)
    if return_smatrix:
        # each block was written column-major into a contiguous n*n slot
        Sg = np.stack([Sg[i].reshape(n, n).T for i in range(4)])
        return R, T, tot, Sg
    return R, T, tot

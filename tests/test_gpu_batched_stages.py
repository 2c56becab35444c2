"""The HBM-resident batched stage entry points (device pointers in and out):
rcwa_layer_smatrix_batched (SMatrixLayer, rcwa.cpp:254-352, with a separate
incidence per layer) and rcwa_star_product_batched (RedhefferProduct,
rcwa.cpp:588-615), against the host-buffer entry points, which run the same
kernels, and against the CPU FP64 oracle."""
import numpy as np
import pytest

import oracle_py as O
import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def _colmajor(mats):
    """[b][n][n] -> contiguous buffer of b column-major blocks."""
    return np.ascontiguousarray(np.stack([m.T for m in mats]))


def _from_colmajor(buf, b, n):
    return np.stack([buf[i].reshape(n, n).T for i in range(b)])


def test_layer_smatrix_batched_vs_host_entry_and_oracle(ctx):
    rng = np.random.default_rng(12)
    probs = [W.random_small(rng, hx=2, hy=1, L=1, grid=16) for _ in range(3)]
    hx, hy = 2, 1
    PQ = (2 * hx + 1) * (2 * hy + 1)
    n = 2 * PQ
    ercs, urcs, Kxs, Kys, k0t = [], [], [], [], []
    for p in probs:   # the same periods and harmonics, different incidences
        p.Lx, p.Ly = probs[0].Lx, probs[0].Ly
        k, Kx, Ky, _, _ = O.wave_vectors(p.Lx, p.Ly, p.lam0, p.theta, p.phi, hx, hy,
                                         p.er_ref, p.ur_ref, p.er_trn, p.ur_trn)
        ercs.append(O.convmat(p.er[0], hx, hy))
        urcs.append(O.convmat(p.ur[0], hx, hy))
        Kxs.append(Kx)
        Kys.append(Ky)
        k0t.append(2 * np.pi / p.lam0 * p.thick[0])
    b = len(probs)
    bufs = [G.DeviceBuffer(b * PQ * PQ * 16) for _ in range(2)] + \
           [G.DeviceBuffer(b * n * n * 16) for _ in range(2)]
    try:
        bufs[0].upload(_colmajor(ercs))
        bufs[1].upload(_colmajor(urcs))
        ctx.layer_smatrix_batched(bufs[0].ptr, bufs[1].ptr, PQ, b, np.stack(Kxs), np.stack(Kys),
                                  k0t, bufs[2].ptr, bufs[3].ptr)
        S11 = _from_colmajor(bufs[2].download((b, n * n)), b, n)
        S12 = _from_colmajor(bufs[3].download((b, n * n)), b, n)
    finally:
        for x in bufs:
            x.free()
    for i in range(b):
        h11, h12 = ctx.layer_smatrix([ercs[i]], [urcs[i]], Kxs[i], Kys[i], [k0t[i]])
        assert np.abs(S11[i] - h11[0]).max() < 1e-13 and np.abs(S12[i] - h12[0]).max() < 1e-13
        r11, r12 = O.smatrix_layer(ercs[i], urcs[i], Kxs[i], Kys[i], 2 * np.pi / probs[i].lam0,
                                   probs[i].thick[0])
        assert np.abs(S11[i] - r11).max() < 1e-10 and np.abs(S12[i] - r12).max() < 1e-10


def _blocks(rng, b, n):
    return [0.3 * (rng.standard_normal((b, n, n)) + 1j * rng.standard_normal((b, n, n)))
            for _ in range(4)]


def test_star_product_batched_vs_host_entry_and_oracle(ctx):
    rng = np.random.default_rng(9)
    b, n = 3, 18
    A, B = _blocks(rng, b, n), _blocks(rng, b, n)
    dA = [G.DeviceBuffer(b * n * n * 16) for _ in range(4)]
    dB = [G.DeviceBuffer(b * n * n * 16) for _ in range(4)]
    dS = [G.DeviceBuffer(b * n * n * 16) for _ in range(4)]
    try:
        for k in range(4):
            dA[k].upload(_colmajor(A[k]))
            dB[k].upload(_colmajor(B[k]))
        ctx.star_batched(n, b, [x.ptr for x in dA], [x.ptr for x in dB], [x.ptr for x in dS])
        S = [_from_colmajor(dS[k].download((b, n * n)), b, n) for k in range(4)]
        # in place: S = A
        ctx.star_batched(n, b, [x.ptr for x in dA], [x.ptr for x in dB], [x.ptr for x in dA])
        Sa = [_from_colmajor(dA[k].download((b, n * n)), b, n) for k in range(4)]
    finally:
        for x in dA + dB + dS:
            x.free()
    host = ctx.star(tuple(A), tuple(B))
    for k in range(4):
        assert np.abs(S[k] - host[k]).max() < 1e-14
        assert np.array_equal(S[k], Sa[k])
    for i in range(b):
        ref = O.redheffer([x[i] for x in A], [x[i] for x in B])
        for k in range(4):
            assert np.abs(S[k][i] - ref[k]).max() < 1e-12

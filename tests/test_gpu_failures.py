"""Failure paths of the solve (the reference aborts or returns inf/NaN at
these points: rcwa.cpp:300 calls abort() when the QR iteration does not
converge, and rcwa.cpp:225, 329, 541, 560 divide by k_z or lambda without a
check).  Every case is forced from the host, so every kernel still exits
normally, and each test then runs an ordinary solve on the same context to
show that it is still usable.  Also the per-result conditioning bound
rcwa_results.cond, pinned against LAPACK."""
import math

import numpy as np
import pytest

import np_rcwa as NP
import oracle_py as O
import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu

E_ARG, E_NOMEM, E_NOCONV, E_SINGULAR, E_NONFINITE = 1, 2, 3, 4, 5


def _small(seed=3, h=2, L=2):
    return W.random_small(np.random.default_rng(seed), hx=h, hy=h, L=L, grid=16)


def _still_works(ctx):
    p = _small(11)
    R, T, tot = ctx.solve([p])[0]
    Ro, To, toto = O.solve(p)
    assert np.abs(R - Ro).max() < 1e-9 and np.abs(T - To).max() < 1e-9


def _code(ctx, p):
    with pytest.raises(G.RcwaError) as e:
        ctx.solve([p])
    return e.value.code


def test_qr_sweep_limit_gives_noconv(ctx):
    """A QR sweep limit of 0 per row stops every matrix at its first sweep:
    RCWA_E_EIG_NOCONV instead of the reference's abort()."""
    ctx.set_eig_maxit(0)
    try:
        assert _code(ctx, _small()) == E_NOCONV
    finally:
        ctx.set_eig_maxit(30)
    _still_works(ctx)


def test_nan_grid_gives_nonfinite(ctx):
    p = _small()
    p.er = p.er.copy()
    p.er[1, 3, 5] = complex(math.nan, 0.0)
    assert _code(ctx, p) == E_NONFINITE
    p = _small()
    p.ur = p.ur.copy()
    p.ur[0, 0, 0] = complex(0.0, math.inf)
    assert _code(ctx, p) == E_NONFINITE
    _still_works(ctx)


def test_nonfinite_scalars(ctx):
    p = _small()
    p.thick = np.array([50.0, math.nan])
    assert _code(ctx, p) == E_NONFINITE
    p = _small()
    p.lam0 = math.inf
    assert _code(ctx, p) == E_NONFINITE
    p = _small()
    p.er_trn = complex(2.0, math.nan)
    assert _code(ctx, p) == E_NONFINITE
    p = _small()
    p.thick = np.array([50.0, -1.0])
    assert _code(ctx, p) == E_ARG
    _still_works(ctx)


def test_over_budget_gives_nomem():
    p = W.c1_single()
    with G.Context(0, hbm_budget=1 << 20) as small:
        with pytest.raises(G.RcwaError) as e:
            small.solve([p])
        assert e.value.code == E_NOMEM


def test_wood_anomaly_gives_singular(ctx):
    """lambda = Lx at normal incidence puts the (+-1, 0) orders at grazing in
    air: k_z = 0 exactly in the superstrate and the gap medium."""
    rng = np.random.default_rng(2)
    p = W.random_small(rng, hx=1, hy=1, L=1, grid=16)
    p.Lx = p.Ly = 500.0
    p.lam0, p.theta, p.phi = 500.0, 0.0, 0.0
    with pytest.raises(G.RcwaError) as e:
        ctx.solve([p])
    assert e.value.code == E_SINGULAR and "k_z = 0" in str(e.value)
    # the same structure a hair away from the anomaly solves
    p.lam0 = 500.0 * (1 - 1e-9)
    ctx.solve([p])
    _still_works(ctx)


def _cond_lapack(p):
    """max over layers of sqrt(n) ||W^-1||_F with W the unit-column
    eigenvectors of Omega^2 = P Q from LAPACK zgeev (column order and phase
    do not change the norm, so the value is the GPU path's as well)."""
    k_inc, Kx, Ky, Kz_ref, Kz_trn = NP.wave_vectors(p.Lx, p.Ly, p.lam0, p.theta, p.phi, p.hx,
                                                    p.hy, p.er_ref, p.ur_ref, p.er_trn, p.ur_trn)
    KX, KY = np.diag(Kx), np.diag(Ky)
    best = 0.0
    for l in range(p.L):
        erc = NP.convmat(p.er[l], p.hx, p.hy)
        urc = NP.convmat(p.ur[l], p.hx, p.hy)
        e_x, e_y = np.linalg.solve(erc, KX), np.linalg.solve(erc, KY)
        u_x, u_y = np.linalg.solve(urc, KX), np.linalg.solve(urc, KY)
        Pm = NP._blk(KX @ e_y, urc - KX @ e_x, KY @ e_y - urc, -KY @ e_x)
        Qm = NP._blk(KX @ u_y, erc - KX @ u_x, KY @ u_y - erc, -KY @ u_x)
        _, Wm = np.linalg.eig(Pm @ Qm)
        n = Wm.shape[0]
        best = max(best, math.sqrt(n) * np.linalg.norm(np.linalg.inv(Wm)))
    return best


def test_cond_matches_lapack(ctx):
    for seed in (1, 4):
        p = _small(seed, h=2, L=3)
        ctx.solve([p])
        got = ctx.last_cond[0]
        want = _cond_lapack(p)
        assert np.all(got >= 1.0 - 1e-12)
        assert np.allclose(got, want, rtol=1e-6), (got, want)
    # no layers: no mode matrices
    p = _small(5, h=1, L=1)
    p.er, p.ur, p.thick = p.er[:0], p.ur[:0], p.thick[:0]
    ctx.solve([p])
    assert np.all(ctx.last_cond[0] == 0.0)

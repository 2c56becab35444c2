"""Serving form: requests of different harmonic counts, layer counts and
wavelengths, submitted while others run, are batched together; each result
is the same alone (plain per-call path) and in the mix."""
import time

import numpy as np
import pytest

import oracle_py as O
import rcwa_amd as G
import workloads as W

pytestmark = pytest.mark.gpu


def _mix(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        h = int(rng.choice([1, 2, 3]))
        L = int(rng.integers(1, 5))
        p = W.random_small(rng, hx=h, hy=h, L=L, grid=16)
        out.append(p)
    return out


def test_mixed_stream_same_alone_and_in_mix(ctx):
    probs = _mix(21, 24)
    with G.Server(0, hbm_budget=8 << 30, n_workers=2) as srv:
        tickets = []
        for i, p in enumerate(probs):
            tickets.append(srv.submit(p))
            if i % 6 == 5:
                time.sleep(0.02)          # later requests arrive while earlier ones run
        res = [srv.wait(t) for t in tickets]
    batched = [r[3]["batch_size"] for r in res]
    assert max(batched) > 1, "no request was batched with another"
    for p, r in zip(probs, res):
        alone = ctx.solve([p])[0]
        assert np.abs(r[0] - alone[0]).max() < 1e-10
        assert np.abs(r[1] - alone[1]).max() < 1e-10
        assert r[3]["t_finish"] >= r[3]["t_start"] >= r[3]["t_submit"]
    for p, r in list(zip(probs, res))[:3]:
        R, T, tot = O.solve(p)
        assert np.abs(r[0] - R).max() < 1e-9 and np.abs(r[1] - T).max() < 1e-9


def test_cancel_and_poll():
    probs = _mix(5, 6)
    with G.Server(0, hbm_budget=4 << 30, n_workers=1, max_batch=1) as srv:
        ts = [srv.submit(p) for p in probs]
        cancelled = 0
        try:
            srv.cancel(ts[-1])
            cancelled = 1
        except G.RcwaError:
            pass                                 # already started: nothing to cancel
        for t in ts[:-1]:
            assert srv.wait(t) is not None
        if cancelled:
            with pytest.raises(G.RcwaError):
                srv.poll(ts[-1])


def test_cancel_wakes_a_waiter_and_tickets_are_collected_once():
    """A thread blocked in wait on a queued ticket returns CANCELLED when
    another thread cancels it (it must not hang), and a collected ticket is
    gone: a second poll reports an unknown ticket."""
    import threading
    probs = _mix(7, 8)
    with G.Server(0, hbm_budget=4 << 30, n_workers=1, max_batch=1) as srv:
        ts = [srv.submit(p) for p in probs]
        got = {}

        def waiter():
            try:
                got["r"] = srv.wait(ts[-1])
            except G.RcwaError as e:
                got["e"] = e.code

        th = threading.Thread(target=waiter)
        th.start()
        time.sleep(0.05)
        try:
            srv.cancel(ts[-1])
            cancelled = True
        except G.RcwaError:
            cancelled = False                    # already taken into a batch
        th.join(timeout=60)
        assert not th.is_alive(), "wait did not return"
        if cancelled:
            assert got.get("e") == 8             # RCWA_E_CANCELLED
        else:
            assert "r" in got
        assert srv.wait(ts[0]) is not None
        with pytest.raises(G.RcwaError):
            srv.poll(ts[0])                      # already collected
        for t in ts[1:-1]:
            srv.wait(t)


@pytest.mark.parametrize("max_batch", [1, 0])
def test_request_of_several_incidences(ctx, max_batch):
    """A request whose sources name three incidences (three wavelengths, one
    at another angle) runs as three flights -- in parts, one after another,
    on a server of one slot (max_batch=1), side by side on the default one --
    and each source's result is the one the single-incidence solve gives."""
    import copy
    rng = np.random.default_rng(31)
    p = W.random_small(rng, hx=2, hy=2, L=3, grid=16)
    incid = [(p.lam0, p.theta, p.phi), (1.3 * p.lam0, p.theta, p.phi),
             (0.8 * p.lam0, 0.5 * p.theta, p.phi + 0.3)]
    sources = [(lam, th, ph, a, b) for lam, th, ph in incid for a, b in p.pols]
    with G.Server(0, hbm_budget=4 << 30, n_workers=1, max_batch=max_batch) as srv:
        t = srv.submit(p, sources=sources)
        R, T, tot, info = srv.wait(t)
    npol = len(p.pols)
    for k, (lam, th, ph) in enumerate(incid):
        q = copy.copy(p)
        q.lam0, q.theta, q.phi = lam, th, ph
        Ra, Ta, tota = ctx.solve([q])[0]
        sl = slice(k * npol, (k + 1) * npol)
        assert np.abs(R[sl] - Ra).max() < 1e-10
        assert np.abs(T[sl] - Ta).max() < 1e-10
        assert np.abs(tot[sl] - tota).max() < 1e-10

"""Writes the full-size golden fixtures of the RCWA solve path from the CPU
FP64 oracle (oracle/liboracle.so, built by `make -C oracle`).

Each fixture is a small .npz of inputs-by-name and expected outputs: the
per-order reflected and transmitted efficiencies R, T ([pol][Q][P], the
Q x P layout of rcwa.cpp:576-582) and the totals, for one problem that the
GPU tests rebuild from rcwa-cuda_amd/workloads.py with the same arguments.

  c2_lam<i>.npz   the 21x21-harmonic, 20-layer trapezoid line grating on
                  512 x 512 grids at wavelength index i of the 200-point
                  250-800 nm sweep, theta 65 deg, TE and TM (index 14,
                  288.69 nm, is the substrate cutoff of the (+1, 0) order,
                  the sweep's worst-conditioned point)
  c5_2layers.npz  the first two layers of the 41x41-harmonic hole-array
                  stack on 1024 x 1024 grids, lambda 400 nm, theta 30 deg

The oracle is single-threaded and naive-loop: a C2 wavelength takes about
a quarter of an hour of one core, the C5 pair of layers about four hours
(13,829 s measured).
Run once, in the build container:

    python tests/golden/make_fixtures.py c2 14 60 120 199   # one process each
    python tests/golden/make_fixtures.py c5
"""
import os
import sys
import time
from multiprocessing import Process

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("oracle", "rcwa-cuda_amd"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import oracle_py  # noqa: E402
import workloads  # noqa: E402


def _save(name, prob, R, T, tot, seconds, **extra):
    path = os.path.join(HERE, name)
    np.savez(path, R=R, T=T, tot=tot, lam0=prob.lam0, theta=prob.theta, phi=prob.phi,
             hx=prob.hx, hy=prob.hy, L=prob.L, pols=np.array(prob.pols),
             oracle_seconds=seconds, **extra)
    print("wrote %s (%.0f s)" % (path, seconds), flush=True)


def c2(index):
    prob = workloads.c2_sweep(indices=[index])[0]
    t0 = time.time()
    R, T, tot = oracle_py.solve(prob)
    _save("c2_lam%d.npz" % index, prob, R, T, tot, time.time() - t0, index=index)


def c5():
    full = workloads.c5_large()
    prob = workloads.Problem(full.er[:2], full.ur[:2], full.thick[:2], full.Lx, full.Ly,
                             full.lam0, full.theta, full.phi, full.pols, full.hx, full.hy,
                             full.er_ref, full.ur_ref, full.er_trn, full.ur_trn)
    del full
    t0 = time.time()
    R, T, tot = oracle_py.solve(prob)
    _save("c5_2layers.npz", prob, R, T, tot, time.time() - t0)


def main(argv):
    if not argv:
        print(__doc__)
        return 2
    if argv[0] == "c2":
        procs = [Process(target=c2, args=(int(i),)) for i in argv[1:]]
    elif argv[0] == "c5":
        procs = [Process(target=c5)]
    else:
        print(__doc__)
        return 2
    for p in procs:
        p.start()
    for p in procs:
        p.join()
    return max(p.exitcode for p in procs)


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))

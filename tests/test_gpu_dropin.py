"""The rcwa.hpp drop-in (include/rcwa_amd.hpp): a caller written like
rcwa3d/src/main.cpp:112-198 runs the stage functions one by one on the GPU
and reproduces the oracle, as does the one-call Solve().  The same caller
then uses every helper of rcwa.hpp -- Vec_3d (rcwa.hpp:17, the polarisation
of rcwa.cpp:481-498), unsigned_sqrt over a matrix (rcwa.hpp:37-44),
MeshGrid (rcwa.hpp:218-224), ComplexEigenSolver + extraction (rcwa.hpp:227-233,
rcwa.cpp:617-621), PartialPivLU and the matrix product -- each checked
against what it must equal."""
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
import workloads as W

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "main_dropin")


def _stdin(p, pol):
    L, Ny, Nx = p.er.shape
    med = []
    for z in (p.er_ref, p.ur_ref, p.er_trn, p.ur_trn):
        med += [z.real, z.imag]
    ints = [Nx, Ny, L]
    flts = [p.Lx, p.Ly]
    ints2 = [p.hx, p.hy]
    flts2 = [p.lam0, p.theta, p.phi, *p.pols[pol]] + med
    body = ([str(v) for v in ints] + [repr(float(v)) for v in flts] + [str(v) for v in ints2]
            + [repr(float(v)) for v in flts2] + [repr(float(t)) for t in p.thick])
    for g in (p.er, p.ur):
        f = g.reshape(-1)
        body += ["%r %r" % (float(v.real), float(v.imag)) for v in f]
    return "\n".join(body) + "\n"


@pytest.mark.parametrize("seed,hx,hy,L", [(0, 1, 1, 2), (1, 2, 1, 3)])
def test_dropin_main_matches_oracle(seed, hx, hy, L):
    assert os.path.exists(EXE), "tests/dropin/main_dropin is not built"
    rng = np.random.default_rng(seed)
    p = W.random_small(rng, hx=hx, hy=hy, L=L, grid=8)
    for pol in range(len(p.pols)):
        out = subprocess.run([EXE], input=_stdin(p, pol), capture_output=True, text=True,
                             timeout=120)
        assert out.returncode == 0, out.stderr
        vals = [list(map(float, l.split())) for l in out.stdout.strip().splitlines()
                if not l.startswith("helper ")]
        PQ = int(vals[0][0])
        staged = np.array(vals[1:1 + PQ])
        tot_s = vals[1 + PQ]
        one = np.array(vals[2 + PQ:2 + 2 * PQ])
        tot_o = vals[2 + 2 * PQ]
        R, T, tot = O.solve(p)
        Rf = R[pol].reshape(-1, order="F")
        Tf = T[pol].reshape(-1, order="F")
        helpers = {l.split()[1]: float(l.split()[2])
                   for l in out.stdout.strip().splitlines() if l.startswith("helper ")}
        assert set(helpers) == {"vec3d", "usqrt", "meshgrid", "eig", "lu", "gemm"}, helpers
        assert helpers["vec3d"] < 1e-14 and helpers["usqrt"] == 0 and helpers["meshgrid"] == 0
        assert helpers["eig"] < 1e-12 and helpers["lu"] < 1e-13 and helpers["gemm"] < 1e-14
        for arr, tt in ((staged, tot_s), (one, tot_o)):
            assert np.abs(arr[:, 0] - Rf).max() < 1e-9
            assert np.abs(arr[:, 1] - Tf).max() < 1e-9
            assert abs(tt[0] - tot[pol][0]) < 1e-9 and abs(tt[1] - tot[pol][1]) < 1e-9

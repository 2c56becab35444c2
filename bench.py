"""Benchmark of the MI355X RCWA solve path on the 21x21-harmonic, 20-layer
spectral sweep (1-D trapezoid line grating staircased into 20 slices over 5
dispersive materials, 512x512 grids, lambda 250-800 nm in 200 points,
theta = 65 deg, TE + TM sharing one S-matrix per wavelength).

One step = one batched pass of the plain FP64 path over `--lams` wavelengths
of the sweep (HBM-resident permittivity grids -> per-order R/T on the host).
The JSON line reports:
  value        solves/s (one solve = one wavelength with both polarisations)
  latency_ms   p50/p99 of the per-step wall time (every solve of a step
               completes with the step)
  roofline     the QR phase of the eigensolver (multishift sweep kernel with
               aggressive early deflation, log apply to the Schur vectors,
               the two products forming the Schur form; the dominant phase),
               classed as HBM-bound (no matrix-core work): its compulsory
               bytes per launch over its event-timed average duration,
               against the HBM peak, beside the traffic measured with PMC
               counters (about 111x the compulsory bytes, moved at 3.9 TB/s:
               the phase is bound by the latency of its one-wave steps, not
               by HBM) and its FP64 flop rate
  split_sample the split-precision form on 8 of the timed wavelengths:
               max |dR|, |dT| against the FP64 results of this run
  cpu_baseline the CPU FP64 oracle on a bounded sample of the same
               workload, run after the GPU phases as one single-threaded
               worker process per host core of the run's CPU share (half
               time one layer S-matrix, half one Redheffer product),
               extrapolated to a full solve: per-core and all-core solves/s
  split        (--split) the split-precision form on the same wavelengths:
               solves/s and max |dR|, |dT| against this FP64 run
  planner      (--planner) the same wavelengths through a sweep plan (the
               grating as material masks whose convolution matrices are formed
               once), in chunks of --lams: solves/s, plan set-up time and max
               |dR|, |dT| against this FP64 run
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rcwa-cuda_amd"))

METRIC = "RCWA solves/s at 21×21 orders, 20 layers; p50/p99 latency; max |ΔR| vs FP64"
# HBM bytes (read + written) of the QR phase per 882x882 layer matrix, from
# rocprofv3 FETCH_SIZE and WRITE_SIZE passes over the QR phase's kernels at
# 510 matrices per launch (the mean of the step's two launches, 3513.0 and
# 3523.8 GB), scaled by the counters' reading on a copy of known size
# (--hbm-copy-probe; profiles/qr_phase_21x21_pmc_traffic.md).
QR_PHASE_HBM_BYTES_PER_MATRIX = 3518.4e9 / 510
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBPS = 8000.0


def cpu_workers():
    """Host cores of this run's CPU share: the affinity mask, capped at the
    16 cores a one-GPU box provides."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def cpu_baseline():
    """One single-threaded oracle process per core of the CPU share, all at
    once: even workers time one SMatrixLayer, odd ones one RedhefferProduct
    (N = 441, the sweep's size).  A solve is 20 layers + 22 star products."""
    W = cpu_workers()
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1")
    kinds = ["layer" if k % 2 == 0 else "star" for k in range(W)] if W > 1 else ["both"]
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-sample", k],
                              stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, text=True)
             for k in kinds]
    t = {"layer": [], "star": []}
    for pr in procs:
        o, e = pr.communicate(timeout=1800)
        if pr.returncode != 0:
            raise RuntimeError("CPU baseline worker failed: " + e[-500:])
        for line in o.strip().splitlines():
            d = json.loads(line)
            t[d["which"]].append(d["seconds"])
    tl, ts = float(np.mean(t["layer"])), float(np.mean(t["star"]))
    per_solve = 20 * tl + 22 * ts
    return {"value": W / per_solve, "unit": "solves/s", "cores": W, "kind": "port",
            "per_core_solves_per_s": 1.0 / per_solve, "cpu_model": cpu_model(),
            "sample": "CPU FP64 oracle (naive-loop restatement of the reference's Eigen path), "
                      "%d single-threaded workers at once on %d cores: %d timed one SMatrixLayer "
                      "(mean %.1f s), %d one RedhefferProduct (mean %.1f s) at N=441; "
                      "extrapolated to 20 layers + 22 star products per solve, x %d cores"
                      % (W, W, len(t["layer"]), tl, len(t["star"]), ts, W)}


def cpu_sample(which):
    """Times one oracle stage at the sweep's size (runs in a child process)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    import workloads as W
    p = W.c2_sweep(indices=[0])[0]
    k, Kx, Ky, Kzr, Kzt = O.wave_vectors(p.Lx, p.Ly, p.lam0, p.theta, p.phi, p.hx, p.hy,
                                         p.er_ref, p.ur_ref, p.er_trn, p.ur_trn)
    erc = O.convmat(p.er[10], p.hx, p.hy)
    urc = O.convmat(p.ur[10], p.hx, p.hy)
    t0 = time.perf_counter()
    S11, S12 = O.smatrix_layer(erc, urc, Kx, Ky, 2 * math.pi / p.lam0, p.thick[10])
    t1 = time.perf_counter()
    if which in ("layer", "both"):
        print(json.dumps({"which": "layer", "seconds": t1 - t0}), flush=True)
    if which in ("star", "both"):
        S = (S11, S12, S12, S11)
        t0 = time.perf_counter()
        O.redheffer(S, S)
        t1 = time.perf_counter()
        print(json.dumps({"which": "star", "seconds": t1 - t0}), flush=True)


class Hip:
    """Minimal HIP runtime access for HBM-resident inputs."""

    def __init__(self):
        self.lib = ctypes.CDLL("libamdhip64.so")

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        rc = self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError("hipMalloc failed: %d" % rc)
        return p.value

    def upload(self, dst, arr):
        rc = self.lib.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(arr.ctypes.data),
                                ctypes.c_size_t(arr.nbytes), 1)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed: %d" % rc)

    def copy(self, dst, src, nbytes):
        rc = self.lib.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src),
                                ctypes.c_size_t(nbytes), 3)  # device to device
        if rc != 0:
            raise RuntimeError("hipMemcpy failed: %d" % rc)

    def sync(self):
        self.lib.hipDeviceSynchronize()

    def free(self, p):
        self.lib.hipFree(ctypes.c_void_p(p))


def serve_bench(args):
    """Serving form on a prefix of the seeded Poisson stream (mean 2000
    req/s, harmonics {7,11,15,21,25}^2, 1-30 layers, 256x256 grids): each
    request is submitted at its arrival time.  The plain per-call path's rate
    is measured on the same requests, run one per call, which also checks
    that each of them gives the same result alone as in the mix."""
    import threading
    import rcwa_amd as G
    import workloads as W
    stream = W.c4_stream(count=args.serve)
    if args.autotune:
        # warm-up pass: the shape classes the stream meets are tuned here,
        # outside the measured pass (the tuned table is process-wide)
        with G.Server(0, n_workers=2, autotune=True) as srv:
            for t in [srv.submit(p) for _, p in stream]:
                srv.wait(t)
        print("serving warm-up pass done (%d tuned shape classes)" % len(G.autotune_entries()),
              file=sys.stderr, flush=True)
    with G.Server(0, n_workers=2, autotune=args.autotune) as srv:
        tickets = [None] * len(stream)
        t0 = time.perf_counter()

        def submitter():
            for i, (ta, p) in enumerate(stream):
                d = t0 + ta - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                tickets[i] = srv.submit(p)
        th = threading.Thread(target=submitter)
        c0 = srv.clock()
        th.start()
        th.join()
        res = []
        for i, t in enumerate(tickets):
            res.append(srv.wait(t))
            if (i + 1) % 50 == 0:
                print("served %d/%d" % (i + 1, len(tickets)), file=sys.stderr, flush=True)
    lat = np.array([r[3]["t_finish"] - r[3]["t_submit"] for r in res])
    first = min(r[3]["t_submit"] for r in res)
    last = max(r[3]["t_finish"] for r in res)
    served = len(res) / (last - first)
    nsample = len(stream)
    ctx = G.Context(0)
    ctx.set_autotune(args.autotune)
    t1 = time.perf_counter()
    alone = []
    for i, (_, p) in enumerate(stream[:nsample]):
        alone.append(ctx.solve([p])[0])
        if (i + 1) % 8 == 0:
            print("per-call %d/%d" % (i + 1, nsample), file=sys.stderr, flush=True)
    percall = nsample / (time.perf_counter() - t1)
    ctx.close()
    dmax = max(float(np.abs(a[0] - r[0]).max()) for a, r in zip(alone, res))
    return {"requests": len(res), "percall_sample": nsample,
            "solves_per_s": served, "percall_solves_per_s": percall,
            "speedup_vs_percall": served / percall,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3),
                           "p99": float(np.percentile(lat, 99) * 1e3)},
            "mean_batch": float(np.mean([r[3]["batch_size"] for r in res])),
            "tuned_shape_classes": len(G.autotune_entries()),
            "tuned_nondefault": sum(1 for e in G.autotune_entries() if e[5] != 0),
            "max_dR_mix_vs_alone": dmax, "clock_start": c0}


def c5_bench():
    """The large single structure: 41x41 harmonics (n = 3362), 50 layers on
    1024^2 grids rasterised on the device from its shape list, one solve.
    Reports the wall time (rasterisation included), the stage times and the
    algorithmic rate."""
    import rcwa_amd as G
    import workloads as W
    Lx, Ly, bg, shapes = W.c5_shapes()
    p = W.c5_large(grid=8)  # the structure's metadata; its grids come from the shapes
    N = 41 * 41
    os.environ["RCWA_PROGRESS"] = "1"  # one line per layer batch: the solve takes a while
    ctx = G.Context(0)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    er = ctx.rasterize(1024, 1024, Lx, Ly, bg, shapes)
    ur = ctx.rasterize(1024, 1024, Lx, Ly, [1.0] * len(bg), [])
    r = ctx.solve([p], device_grids=[(er.ptr, ur.ptr)], grid_shape=(1024, 1024))[0]
    wall = time.perf_counter() - t0
    er.free()
    ur.free()
    del os.environ["RCWA_PROGRESS"]
    st = ctx.stats()
    ctx.close()
    flops = (2384.0 * 50 + 1621.0) * N ** 3
    stages = sum(st[k] for k in ("ms_convmat", "ms_layers", "ms_fold", "ms_out")) * 1e-3
    return {"workload": "41x41 harmonics (N=1681, 3362x3362 complex128), 50 layers, 1024x1024 "
                        "grids, lambda 400 nm, theta 30 deg, TE; grids rasterised on the device",
            "seconds": wall, "tflops_algorithmic": flops / wall / 1e12,
            # the solve's own stages (HIP events); the wall time adds the
            # rasterisation and the context's first HBM allocation, which
            # has taken up to 5 s longer on some runs than on others
            "stage_seconds": stages,
            "R_tot": float(r[2][0][0]), "T_tot": float(r[2][0][1]),
            "stage_ms": {k: st[k] for k in ("ms_convmat", "ms_layers", "ms_eig", "ms_eig_hess",
                                            "ms_eig_schur", "ms_eig_vec", "ms_fold")}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    # 51 wavelengths x 20 layers = 1020 layer eigenproblems, solved in two
    # launches of 510: about two full rounds of one matrix per CU on 256 CUs
    # (48 gives 480 per launch, with 32 CUs idle in the second round)
    ap.add_argument("--lams", type=int, default=51, help="wavelengths per step")
    ap.add_argument("--split", action="store_true", help="also time the split-precision form")
    ap.add_argument("--planner", action="store_true", help="also time the sweep-plan form")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-split-sample", action="store_true",
                    help="leave out the split-precision accuracy sample (outside the timed "
                         "region), e.g. for counter passes over the FP64 path alone")
    ap.add_argument("--refine", type=int, default=None,
                    help="fold refinement level (rcwa_ctx_set_refine; default: the library's)")
    ap.add_argument("--c5", action="store_true",
                    help="also time the 41x41-harmonic, 50-layer single structure")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the complex GEMM tile shapes on first use (rcwa_ctx_set_autotune; "
                         "the warm-up step pays the tuning)")
    ap.add_argument("--serve", type=int, default=0,
                    help="also run the serving form on this many requests of the stream")
    ap.add_argument("--cpu-sample", choices=["layer", "star", "both"], help=argparse.SUPPRESS)
    ap.add_argument("--hbm-copy-probe", action="store_true",
                    help="only copy 1 GiB between two device buffers three times (the runtime's "
                         "copy kernel): the known-size reference for HBM traffic counters")
    args = ap.parse_args()
    if args.cpu_sample:
        cpu_sample(args.cpu_sample)
        return
    if args.hbm_copy_probe:
        hip = Hip()
        a, b = hip.malloc(1 << 30), hip.malloc(1 << 30)
        for _ in range(3):
            hip.copy(b, a, 1 << 30)
        hip.sync()
        hip.free(a)
        hip.free(b)
        print(json.dumps({"hbm_copy_probe_bytes": 1 << 30, "copies": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the multi-GPU form of this path is design-only "
              "(wavelengths/requests shard with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import rcwa_amd as G
    import workloads as W

    nlam_needed = (args.warmup + args.steps) * args.lams

    # wavelengths spread over the 200-point sweep: warmup chunks first, then the
    # timed chunks; a run of more than 200 solves revisits wavelengths.  One
    # device grid per distinct wavelength.
    idx = np.linspace(0, 199, nlam_needed).round().astype(int).tolist()
    hip = Hip()
    ur_dev = None
    one = {}
    for i in sorted(set(idx)):
        p = W.c2_sweep(indices=[i])[0]
        if ur_dev is None:
            ur_dev = hip.malloc(p.ur.nbytes)
            hip.upload(ur_dev, p.ur)
        d = hip.malloc(p.er.nbytes)
        hip.upload(d, p.er)
        p.er = p.er[:, :1, :1]  # host copies no longer needed; keep shape metadata small
        p.ur = p.ur[:, :1, :1]
        one[i] = (p, d)
    probs = [one[i][0] for i in idx]
    er_devs = [one[i][1] for i in idx]
    hip.sync()

    ctx = G.Context(0)
    ctx.set_autotune(args.autotune)
    if args.refine is not None:
        ctx.set_refine(args.refine)

    def run(chunk, precision=G.RCWA_FP64):
        sel = probs[chunk * args.lams:(chunk + 1) * args.lams]
        grids = [(er_devs[chunk * args.lams + i], ur_dev) for i in range(len(sel))]
        return ctx.solve(sel, precision, device_grids=grids, grid_shape=(512, 512))

    def progress(msg):
        print("[bench] " + msg, file=sys.stderr, flush=True)

    for w in range(args.warmup):
        run(w)
        progress("warm-up step %d done" % (w + 1))
    hip.sync()
    ctx.reset_stats()
    ctx.set_timing(True)
    step_ms = []
    results = []
    t_start = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        results.append(run(args.warmup + k))
        step_ms.append((time.perf_counter() - t0) * 1e3)
        progress("timed step %d done (%.0f ms)" % (k + 1, step_ms[-1]))
    hip.sync()
    elapsed = time.perf_counter() - t_start
    st = ctx.stats()
    ctx.set_timing(False)
    solves = args.steps * args.lams
    value = solves / elapsed

    N = 441
    n = 2 * N
    # QR sweeps with Schur vectors: (25 - 10/3 - 4/3) n^3 real flops (Golub & Van Loan,
    # Matrix Computations, 4th ed., sec. 7.5.6), x4 for complex arithmetic.
    f_qr = 4.0 * (25.0 - 14.0 / 3.0) * n ** 3
    launches = max(st["n_eig_launches"], 1)
    per_launch_mats = st["n_eig_matrices"] / launches
    avg_ms = st["ms_eig_schur"] / launches
    tflops = f_qr * per_launch_mats / (avg_ms * 1e-3) / 1e12 if avg_ms > 0 else 0.0
    # compulsory HBM bytes per matrix: read the Hessenberg matrix, its Q and
    # the copy of A; write the Schur form T and the Schur vectors Z
    compulsory = 5.0 * n * n * 16
    gbps = compulsory * per_launch_mats / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = QR_PHASE_HBM_BYTES_PER_MATRIX * per_launch_mats
    roofline = {"kernel": "QR phase: schur_ms_kernel + qr_apply_ms_kernel + 2 x zgemm_kernel "
                          "(complex multishift QR with aggressive early deflation on the "
                          "Hessenberg matrix: 12-row deflation windows at this size, up to eight chained "
                          "single-shift sweeps per pass; the logged rotations and window "
                          "blocks applied to the Schur vectors Z; Schur form T = Z^H A Z)",
                "bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": gbps / HBM_PEAK_GBPS, "traffic": traffic,
                "traffic_unit": "bytes per launch (FETCH_SIZE + WRITE_SIZE passes, "
                                "copy-kernel calibrated)",
                "compulsory_bytes_per_launch": compulsory * per_launch_mats,
                "traffic_over_compulsory": traffic / (compulsory * per_launch_mats),
                "traffic_rate_gbps": traffic / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0,
                "fp64_tflops": tflops, "fp64_frac": tflops / FP64_PEAK_TFLOPS,
                "avg_launch_ms": avg_ms, "matrices_per_launch": per_launch_mats}
    whole = st["flops_algorithmic"] / (st["ms_total"] * 1e-3) / 1e12 if st["ms_total"] else 0.0

    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "21x21 harmonics (N=441, 882x882 complex128), 20 layers, "
                               "512x512 grids, lambda 250-800 nm sweep, theta 65 deg, TE+TM",
                   "lams_per_step": args.lams, "solves_timed": solves},
        "wavelength_points_per_s": value,
        "latency_ms": {"p50": float(np.percentile(step_ms, 50)),
                       "p99": float(np.percentile(step_ms, 99)),
                       "what": "per-step wall time; every solve of a step finishes with it"},
        "path_fp64_tflops": whole,
        "path_fp64_frac": whole / FP64_PEAK_TFLOPS,
        "stage_ms": {k: st[k] for k in ("ms_convmat", "ms_layers", "ms_eig", "ms_eig_hess",
                                        "ms_eig_schur", "ms_eig_vec", "ms_fold", "ms_out")},
        "roofline": roofline,
    }

    # the split-precision form on 8 of the timed wavelengths, against this run's
    # FP64 results (the metric's max |dR| vs FP64)
    flat = [r for step in results for r in step]
    if not args.no_split_sample:
        pick = np.linspace(0, len(flat) - 1, min(8, len(flat))).round().astype(int).tolist()
        sel = [probs[args.warmup * args.lams + i] for i in pick]
        grids = [(er_devs[args.warmup * args.lams + i], ur_dev) for i in pick]
        sout = ctx.solve(sel, G.RCWA_SPLIT, device_grids=grids, grid_shape=(512, 512))
        out["split_sample"] = {
            "wavelengths": len(pick),
            "max_dR_vs_fp64": float(max(np.abs(flat[i][0] - o[0]).max()
                                        for i, o in zip(pick, sout))),
            "max_dT_vs_fp64": float(max(np.abs(flat[i][1] - o[1]).max()
                                        for i, o in zip(pick, sout))),
            "max_dtot_vs_fp64": float(max(np.abs(flat[i][2] - o[2]).max()
                                          for i, o in zip(pick, sout)))}
        progress("split-precision sample done")

    if args.split:
        sres = []
        t0 = time.perf_counter()
        for k in range(args.steps):
            sres.append(run(args.warmup + k, G.RCWA_SPLIT))
            progress("split-precision step %d done" % (k + 1))
        hip.sync()
        selapsed = time.perf_counter() - t0
        dR = max(np.abs(a[0] - b[0]).max() for ra, rb in zip(results, sres) for a, b in zip(ra, rb))
        dT = max(np.abs(a[1] - b[1]).max() for ra, rb in zip(results, sres) for a, b in zip(ra, rb))
        out["split"] = {"solves_per_s": solves / selapsed, "speedup_vs_fp64": elapsed / selapsed,
                        "max_dR_vs_fp64": float(dR), "max_dT_vs_fp64": float(dT)}

    if args.planner:
        material, thick, lams, eps = W.c2_structure()
        tidx = idx[args.warmup * args.lams:]
        th = math.radians(65.0)
        srcs = [[(lams[i], th, 0.0, 1.0, 0.0), (lams[i], th, 0.0, 0.0, 1.0)] for i in tidx]
        t0 = time.perf_counter()
        sw = ctx.sweep(material, thick, 120.0, 120.0, eps.shape[1], 10, 10)
        t_plan = time.perf_counter() - t0
        ctx.set_chunk(args.lams)
        t0 = time.perf_counter()
        pres = sw.run(srcs, eps[tidx])
        pelapsed = time.perf_counter() - t0
        ctx.set_chunk(0)
        sw.close()
        progress("sweep-plan form done")
        flat = [r for step in results for r in step]
        dR = max(np.abs(a[0] - b[0]).max() for a, b in zip(flat, pres))
        dT = max(np.abs(a[1] - b[1]).max() for a, b in zip(flat, pres))
        out["planner"] = {"solves_per_s": len(tidx) / pelapsed,
                          "speedup_vs_fp64": elapsed / pelapsed,
                          "plan_setup_ms": t_plan * 1e3,
                          "max_dR_vs_fp64": float(dR), "max_dT_vs_fp64": float(dT)}

    # the stream and the large structure below use contexts of their own:
    # release this one's HBM arena and the sweep's grids first
    ctx.close()
    for _, d in one.values():
        hip.free(d)
    hip.free(ur_dev)
    er_devs = []

    if args.serve:
        progress("serving stream of %d requests" % args.serve)
        out["serving"] = serve_bench(args)

    if args.c5:
        progress("41x41-harmonic single structure")
        out["c5"] = c5_bench()

    if not args.no_cpu_baseline:
        progress("CPU baseline: %d oracle workers" % cpu_workers())
        out["cpu_baseline"] = cpu_baseline()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the MI355X neutronics hot path (one JSON line on stdout).

Workload (one "step" = one cold k-eigenvalue solve through the drop-in path:
device assembly of R and F, Wielandt-shifted power iteration with GMRES +
cell-block-Jacobi inner solves, power normalisation):
  2-D quarter core, 170 x 170 cells of 1 cm, 17 x 17 blocks of 4 seeded
  materials, 2 groups, reflective x=0/y=0, vacuum outer edges, eigen residual 1e-8.
`value` is the mixed-precision rate (FP32 inner solves, FP64 outer quantities);
the plain FP64 drop-in rate, p50/p99 latencies and the k (pcm) / scalar-flux
(relative L2) errors of both against the FP64 oracle are reported beside it.
Also reported: the transport sweep of the hexagonal S6 lattice (317,760
prisms, 4 groups, 61 M angular unknowns) in FP32 and FP64, and the SpMV of the
quarter-core operator, each against the 8 TB/s HBM peak.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
# HBM bytes per gmres_spmv_dot launch (all step launches of the ticks, idle ones included)
# on the quarter-core workload: two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE) over
# `bench.py --roofline-pass`, each scaled by what the same counter reads for
# stream_copy_kernel moving 2^28 bytes each way (fetch x2.000, write x1.000):
# 2541.1 KiB x 2.000 + 203.4 KiB = 5.412 MB.
SPMV_DOT_TRAFFIC_BYTES = 5.412e6
# the same counter passes over `bench.py --batched-pass` (256 instances in one mixed pool),
# HBM bytes per launch of each GMRES step kernel
BATCHED_TRAFFIC_BYTES = {"gmres_spmv_dot": 543.1e6, "gmres_orth": 745.2e6, "gmres_poly": 874.1e6}
METRIC = ("eigen solves/s + p50/p99 latency, mixed vs FP64 at |dk|<=1pcm; sweep/SpMV HBM GB/s")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_solves(pa, d, precision, steps, warmup):
    """Cold eigen solves through the pampa_* C-ABI; returns per-solve seconds,
    keff and the power-normalised flux of the last solve."""
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pa.DropIn("input.pmp", "-silent", "-pk_precision", precision) as run:
            C, G, _ = run.sizes()
            zero = np.zeros(C * G)
            lat = []
            for it in range(warmup + steps):
                run.set_field("scalar-flux", zero)  # cold start: no initial space
                check = pa.lib().pk_synchronize()
                assert check == 0
                t0 = time.perf_counter()
                run.solve()
                pa.lib().pk_synchronize()
                t1 = time.perf_counter()
                if it >= warmup:
                    lat.append(t1 - t0)
            k, _ = run.result()
            phi = run.field("scalar-flux")
            st = run.stats()
            # profiled pass (separate from the timed region): per-kernel device time
            pa.profile(True)
            run.set_field("scalar-flux", zero)
            run.solve()
            prof = {}
            for name in ("gmres_spmv_dot", "gmres_orth", "spmv", "spmv_fp32"):
                n, s, b = pa.profile_get(name)
                if n:
                    prof[name] = dict(launches=n, seconds=s, bytes=b)
            pa.profile(False)
            spmv_s, spmv_b = run.time_kernel(200, fp32=(precision == "mixed"))
        return lat, k, phi, st, prof, (spmv_s, spmv_b)
    finally:
        os.chdir(cwd)


def roofline_pass(pa, d, solves=2):
    """Profiled cold solves only (mixed precision): per-kernel launches, device
    time from HIP events on the library stream, algorithmic bytes."""
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pa.DropIn("input.pmp", "-silent", "-pk_precision", "mixed") as run:
            C, G, _ = run.sizes()
            pa.profile(True)
            for _ in range(solves):
                run.set_field("scalar-flux", np.zeros(C * G))
                run.solve()
            out = {}
            for name in ("gmres_spmv_dot", "gmres_orth"):
                n, sec, byt = pa.profile_get(name)
                out[name] = {"launches": n, "avg_us": sec / max(n, 1) * 1e6,
                             "bytes_per_launch": byt / max(n, 1), "GBs": byt / sec / 1e9 if sec else 0.0}
            pa.profile(False)
        n = 1 << 26  # 256 MB per buffer: the calibration kernel of the traffic counters
        sec = pa.stream_copy(n, 20)
        return {"roofline_pass": out, "solves": solves,
                "stream_copy": {"bytes": 8.0 * n, "us": sec * 1e6, "GBs": 8.0 * n / sec / 1e9}}
    finally:
        os.chdir(cwd)


def cpu_baseline(d, target_seconds=15.0):
    """The oracle (FP64 CPU restatement: reference-form assembly, exact sparse LU
    + restarted Arnoldi shift-invert, power normalisation) timed on a bounded
    sample of the same workload, single thread."""
    import scipy.linalg  # noqa: F401  (load every BLAS the solve uses, so that the limit below covers it)
    import scipy.sparse.linalg  # noqa: F401
    import oracle
    from threadpoolctl import threadpool_limits
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with threadpool_limits(1):
            case = oracle.pmp.read_input("input.pmp")
            times = []
            res = None
            t_start = time.perf_counter()
            while (time.perf_counter() - t_start < target_seconds and len(times) < 20) or len(times) < 2:
                t0 = time.perf_counter()
                res = oracle.solve_diffusion_case(case)
                times.append(time.perf_counter() - t0)
        return times, res
    finally:
        os.chdir(cwd)


def _cpu_worker(args):
    d, seconds = args
    import scipy.linalg  # noqa: F401  (load every BLAS the solve uses, so that the limit below covers it)
    import scipy.sparse.linalg  # noqa: F401
    import oracle
    from threadpoolctl import threadpool_limits
    os.chdir(d)
    with threadpool_limits(1):
        case = oracle.pmp.read_input("input.pmp")
        n, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds or n == 0:
            oracle.solve_diffusion_case(case)
            n += 1
        return n, time.perf_counter() - t0


def cpu_baseline_all_cores(d, seconds=10.0, procs=None):
    """The same oracle solves on every host core at once: `procs` processes
    (the host's CPU share, at most 16), each running cold solves of the same
    workload for `seconds`, one thread each.  Aggregate solves/s of the host."""
    import multiprocessing as mp
    procs = procs or min(16, len(os.sched_getaffinity(0)))
    ctx = mp.get_context("spawn")
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_worker, [(d, seconds)] * procs)
    total = sum(n / t for n, t in res)
    return total, procs, sum(n for n, _ in res)


def sn_solve_numbers(pa, d, reps=3):
    """Cold k-eigenvalue solves of the S6 hexagonal lattice through the drop-in
    C-ABI (sweeps inside Anderson-accelerated source iteration), FP64 and mixed
    (FP32 sweeps, FP64 finish): solves/s, sweeps per solve, and the mixed
    eigenpair against the FP64 one.  (The FP64 path is checked against the
    oracle on smaller lattices by the tests; the oracle's CPU solve of 61 M
    angular unknowns takes minutes.)"""
    cwd = os.getcwd()
    os.chdir(d)
    out, sol = {}, {}
    try:
        for name, prec in (("fp64", "fp64"), ("mixed", "mixed")):
            with pa.DropIn("input.pmp", "-silent", "-pk_precision", prec) as run:
                run.solve()  # records built, buffers allocated
                times = []
                for _ in range(reps):
                    run.cold_start()
                    pa.lib().pk_synchronize()
                    t0 = time.perf_counter()
                    run.solve()
                    pa.lib().pk_synchronize()
                    times.append(time.perf_counter() - t0)
                k, _ = run.result()
                sol[name] = (k, run.field("scalar-flux"))
                st = run.stats()
                out[name] = {"solves_per_s": len(times) / sum(times), "ms_per_solve": 1e3 * sum(times) / len(times),
                             "sweeps": st["inner"], "keff": k}
            log(f"[bench]   S6 lattice solve {name}: {out[name]['ms_per_solve']:.1f} ms, {out[name]['sweeps']} sweeps")
    finally:
        os.chdir(cwd)
    (k64, p64), (k32, p32) = sol["fp64"], sol["mixed"]
    out["mixed_over_fp64"] = out["mixed"]["solves_per_s"] / out["fp64"]["solves_per_s"]
    out["mixed_vs_fp64"] = {"dk_pcm": abs(k32 - k64) * 1e5,
                            "flux_l2": float(np.linalg.norm(p32 - p64) / np.linalg.norm(p64))}
    return out


def sweep_numbers(pa, d):
    cwd = os.getcwd()
    os.chdir(d)
    try:
        with pa.DropIn("input.pmp", "-silent") as run:
            out = {}
            for name, fp32 in (("fp32", True), ("fp64", False)):
                s, b = run.time_kernel(5, fp32=fp32)
                out[name] = dict(ms_per_sweep=s * 1e3, bytes_per_sweep=b, GBs=b / s / 1e9,
                                 frac_hbm=b / s / 1e9 / HBM_PEAK_GBS)
            C, G, M = run.sizes()
            out["angular_unknowns"] = C * G * M
        return out
    finally:
        os.chdir(cwd)


def contraction_numbers(pa, W, tmp, solves=True):
    """The S_N scattering + fission source at 16 groups, a per-material 16 x 16
    contraction per cell, on the vector kernel and on the matrix cores (FP64
    MFMA), on the lattice geometry of the S6 workload (331 columns x 24
    triangles x 40 layers = 317,760 cells) with 16 groups (S2: the source does
    not depend on the directions): device time per call, algorithmic bytes
    (moment read, source written, volume and material) and flops (2 C G^2), and
    one cold mixed eigen solve per contraction."""
    d = W.hex_sn_3d(os.path.join(tmp, "hex-g16"), G=16, order=2)
    cwd = os.getcwd()
    os.chdir(d)
    out = {"workload": "hex lattice 331 columns x 24 triangles x 40 layers, G=16, S2"}
    try:
        for name, mode in (("vector", 1), ("mfma", 2)):
            with pa.DropIn("input.pmp", "-silent", "-pk_precision", "mixed",
                           "-pk_contraction", "vector" if mode == 1 else "mfma") as run:
                C, G, _ = run.sizes()
                r = {}
                for prec, fp32 in (("fp32_out", True), ("fp64_out", False)):
                    s, b = run.time_source(20, fp32=fp32, mode=mode)
                    fl = 2.0 * C * G * G
                    r[prec] = {"us": s * 1e6, "bytes": b, "GBs": b / s / 1e9, "frac_hbm": b / s / 1e9 / HBM_PEAK_GBS,
                               "TFLOPs": fl / s / 1e12}
                if solves:
                    run.solve()  # records built
                    run.cold_start()
                    pa.lib().pk_synchronize()
                    t0 = time.perf_counter()
                    run.solve()
                    pa.lib().pk_synchronize()
                    k, _ = run.result()
                    r["eigen_mixed"] = {"ms": 1e3 * (time.perf_counter() - t0), "keff": k,
                                        "sweeps": run.stats()["inner"]}
                out[name] = r
            log(f"[bench]   source contraction G=16 {name}: {r['fp32_out']['us']:.1f} us "
                f"({r['fp32_out']['GBs']:.0f} GB/s)")
    finally:
        os.chdir(cwd)
    out["mfma_speedup"] = out["vector"]["fp32_out"]["us"] / out["mfma"]["fp32_out"]["us"]
    return out


def _shape(v):
    v = np.asarray(v, dtype=float)
    return v / v.sum()


def _oracle_request(oracle, W, path, seed, rel):
    """FP64 oracle solve of a served request (perturbed tables of a base case)."""
    case = oracle.pmp.read_input(path)
    mesh = oracle.build_mesh(case)
    s = case.solver
    xs = oracle.xs_tables(case.materials, s.G, diffusion=True)
    xs.sigt, xs.sigs, xs.D, xs.nusf = W.perturb_tables(xs.sigt, xs.sigs, xs.D, xs.nusf, seed, rel)
    bct, bcv = oracle.bc_table(mesh, s.bcs)
    R, F = oracle.assemble_diffusion(mesh, xs, bct, bcv)
    return oracle.eigen_sinvert(R, F)


def _oracle_request_sn(oracle, W, path, seed, rel):
    """FP64 oracle solve of a served S_N request: source iteration with sweeps
    on the perturbed tables; returns keff and the moment sum_m w_m psi."""
    case = oracle.pmp.read_input(path)
    mesh = oracle.build_mesh(case)
    s = case.solver
    xs = oracle.xs_tables(case.materials, s.G, diffusion=False)
    xs.sigt, xs.sigs, _, xs.nusf = W.perturb_tables(xs.sigt, xs.sigs, None, xs.nusf, seed, rel)
    k, _, phiw, _ = oracle.sn_sweep_eigen(mesh, xs, s, inner=1, tol_k=1e-12, tol_phi=1e-11)
    return k, phiw


def _oracle_request_job(args):
    path, seed, rel = args
    import scipy.linalg  # noqa: F401  (load every BLAS the solve uses, so that the limit below covers it)
    import scipy.sparse.linalg  # noqa: F401
    import oracle
    from pampa_amd import workloads as W
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        k, x = _oracle_request(oracle, W, path, seed, rel)
    return k, x


FIXTURE = os.path.join(HERE, "tests", "golden", "batched_classes.npz")


def _class_accuracy(W, res, paths, sample, rows):
    """Per size class of the batched family: the served keff and flux against
    the FP64 oracle -- solved here in parallel processes (1 thread each) for the
    classes it finishes in about a minute, from the committed oracle fixture
    (keff, every 64th flux entry) for the four largest."""
    import multiprocessing as mp
    fx = np.load(FIXTURE) if os.path.exists(FIXTURE) else None
    direct, out, skipped = [], [], []
    for r in sample:
        key = f"inst{r}_keff"
        if fx is not None and key in fx.files:
            st = int(fx["stride"])
            k_o, x_o = float(fx[key]), fx[f"inst{r}_phi"]
            xs = np.asarray(res["flux"][r])[::st]
            out.append({"instance": int(r), "rows": int(rows[r]), "source": "fixture (every 64th entry)",
                        "dk_pcm": abs(res["keff"][r] - k_o) * 1e5,
                        "flux_l2": float(np.linalg.norm(_shape(xs) - _shape(x_o)) / np.linalg.norm(_shape(x_o)))})
        elif _class_of(W, r) in BIG_CLASSES:
            skipped.append(int(r))
        else:
            direct.append(r)
    procs = min(16, len(os.sched_getaffinity(0)), max(1, len(direct)))
    with mp.get_context("spawn").Pool(procs) as pool:
        sols = pool.map(_oracle_request_job, [(paths[r], 0, 0.0) for r in direct])
    for r, (k_o, x_o) in zip(direct, sols):
        out.append({"instance": int(r), "rows": int(rows[r]), "source": "oracle",
                    "dk_pcm": abs(res["keff"][r] - k_o) * 1e5,
                    "flux_l2": float(np.linalg.norm(_shape(res["flux"][r]) - _shape(x_o)) /
                                     np.linalg.norm(_shape(x_o)))})
    out.sort(key=lambda a: a["rows"])
    return {"max_dk_pcm": max(a["dk_pcm"] for a in out), "max_flux_l2": max(a["flux_l2"] for a in out),
            "classes": len(out), "per_class": out, "skipped_no_fixture": skipped}


def _accuracy(oracle, W, res, paths, cid, seeds, rel, sample, sn_cases=()):
    out = []
    for r in sample:
        solve = _oracle_request_sn if cid[r] in sn_cases else _oracle_request
        k_o, x_o = solve(oracle, W, paths[cid[r]], int(seeds[r]), rel)
        out.append({"request": int(r), "dk_pcm": abs(res["keff"][r] - k_o) * 1e5,
                    "flux_l2": float(np.linalg.norm(_shape(res["flux"][r]) - _shape(x_o)) /
                                     np.linalg.norm(_shape(x_o)))})
    return {"max_dk_pcm": max(a["dk_pcm"] for a in out), "max_flux_l2": max(a["flux_l2"] for a in out),
            "requests": out}


def _serve(pa, paths, precision, slots, arrival, cid, seeds, rel, warm, keep=(), budget=64e9, sn_workers=1):
    srv = pa.Server(budget_bytes=budget, max_slots=slots, precision=precision)
    srv.sn_workers(sn_workers)
    t0 = time.perf_counter()
    for p in paths:
        srv.add_case(p)
    log(f"[bench]   {len(paths)} cases resident ({time.perf_counter() - t0:.1f} s)")
    srv.progress(30.0)
    if warm:
        # one unperturbed solve per base case first: the stream's warm starts
        srv.run(np.zeros(len(paths)), np.arange(len(paths)), np.zeros(len(paths), np.uint64), rel=0.0,
                warm=False)
    res = srv.run(arrival, cid, seeds, rel=rel, warm=warm, keep_flux=list(keep))
    st = srv.stats()
    log(f"[bench]   {len(arrival)} requests done, makespan {float((np.asarray(arrival) + res['latency']).max()):.2f} s")
    srv.close()
    return res, st


def _family(W, tmp, n):
    """Instances 0..n-1 of the batched family: (paths, rows per instance)."""
    paths, rows = [], []
    for i in range(n):
        d = os.path.join(tmp, f"inst-{i}")
        side, G = W.batched_instance(d, i)
        paths.append(os.path.join(d, "input.pmp"))
        rows.append(side * side * G)
    return paths, np.array(rows)


def batched_spmv(pa, paths, reps=20):
    """One SpMV over the inner operators of every instance at once (all
    resident in one pool), FP32 and FP64: HIP events around `reps`
    back-to-back launches on the library stream."""
    spmv = {}
    for name, prec in (("fp32", pa.PK_PREC_MIXED), ("fp64", pa.PK_PREC_FP64)):
        srv = pa.Server(budget_bytes=64e9, max_slots=len(paths), precision=prec)
        for p in paths:
            srv.add_case(p)
        sec, byt, nrows = srv.time_spmv(reps)
        srv.close()
        spmv[name] = {"kernel": f"bspmv_kernel<{'float' if name == 'fp32' else 'double'}>", "us": sec * 1e6,
                      "bytes": byt, "GBs": byt / sec / 1e9, "frac_hbm": byt / sec / 1e9 / HBM_PEAK_GBS,
                      "rows": nrows}
        log(f"[bench]   batched SpMV {name}: {byt / sec / 1e9:.0f} GB/s")
    return spmv


def step_kernel_profile(pa):
    """Per-kernel launches, device seconds (HIP events recorded by each launch)
    and algorithmic bytes of the GMRES step kernels from the library profiler;
    the dominant kernel's roofline."""
    prof = {}
    for name in ("gmres_spmv_dot", "gmres_orth", "gmres_poly"):
        cnt, sec, byt = pa.profile_get(name)
        if cnt:
            prof[name] = dict(launches=cnt, seconds=sec, bytes=byt, avg_us=sec / cnt * 1e6,
                              bytes_per_launch=byt / cnt, GBs=byt / sec / 1e9)
    roof = None
    if prof:
        name, dv = max(prof.items(), key=lambda kv: kv[1]["seconds"])
        ach = dv["bytes"] / dv["seconds"] / 1e9
        roof = {"bound": "hbm", "kernel": name, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "avg_launch_us": dv["avg_us"],
                "bytes_per_launch": dv["bytes_per_launch"], "launches": dv["launches"]}
    return prof, roof


def batched_profile_traffic(prof, roof):
    """Counter-pass HBM bytes per launch of the batched pool's step kernels."""
    if roof:
        roof["traffic"] = BATCHED_TRAFFIC_BYTES.get(roof["kernel"])
    for name, dv in prof.items():
        dv["traffic_per_launch"] = BATCHED_TRAFFIC_BYTES.get(name)
    return prof, roof


def batched_profile(pa, paths):
    """Profiled pass of the batched family in one mixed pool: the same ticks as
    the timed pass, launched one by one with events (every step launch of the
    captured cycle counted, idle ones included)."""
    n = len(paths)
    log(f"[bench] batched: {n} instances in one mixed pool, profiled pass")
    pa.profile(True)
    _serve(pa, paths, pa.PK_PREC_MIXED, n, np.zeros(n), np.arange(n), np.zeros(n, np.uint64), 0.0, False)
    prof, roof = step_kernel_profile(pa)
    pa.profile(False)
    return batched_profile_traffic(prof, roof)


# size classes whose oracle solve takes minutes here: checked against the committed fixture
BIG_CLASSES = {(128, 16), (256, 8), (192, 16), (256, 16)}


def _class_of(W, i):
    rng = W.SplitMix64(3_000_000 + i)
    side = W.BATCH_SIDES[rng.choice(len(W.BATCH_SIDES))]
    return side, W.BATCH_GROUPS[rng.choice(len(W.BATCH_GROUPS))]


def _class_sample(W, tmp, n, rows):
    """One instance of every (side, groups) class present among the first n
    instances of the batched family (the first of each class)."""
    seen, out = set(), []
    for i in range(n):
        c = _class_of(W, i)
        if c not in seen:
            seen.add(c)
            out.append(i)
    return out


def tuner_report(pa):
    """The autotuner's choice per size class of the pools run so far: slice
    height and SpMV workgroup, the chosen candidate's time and the 64-row /
    256-row default's (FP32 values in mixed pools)."""
    out = {}
    for c in pa.tune_choices():
        side = int(round((c["rows"] / c["block"]) ** 0.5))
        name = f"{side}x{side}x{c['block']}" + ("" if c["precision"] == pa.PK_PREC_MIXED else "-fp64")
        if side * side * c["block"] != c["rows"]:
            name = f"rows{c['rows']}xG{c['block']}"
        best = c["us"][f"{c['slice_height']}x{c['spmv_block']}"]
        out[name] = {"slice_height": c["slice_height"], "spmv_block": c["spmv_block"], "us": round(best, 2),
                     "us_64x256": round(c["us"]["64x256"], 2)}
    return out


def batched_numbers(pa, W, oracle, tmp, n_inst=256, fp64_sample=32):
    """Config-3 family: n_inst independent instances (sides and group counts
    drawn by seed, cross sections perturbed +-5% per instance), all admitted at
    once into one mixed-precision pool; the plain FP64 drop-in (one solve at a
    time) on the first fp64_sample of them, and the pool on the same subset."""
    paths, rows = _family(W, tmp, n_inst)
    n = n_inst
    cid = np.arange(n)
    seeds = np.zeros(n, np.uint64)
    small = _class_sample(W, tmp, n, rows)
    prof, roof = batched_profile(pa, paths)
    log("[bench] batched: timed pass")
    res, st = _serve(pa, paths, pa.PK_PREC_MIXED, n, np.zeros(n), cid, seeds, 0.0, False, keep=small)
    spmv = batched_spmv(pa, paths)
    mixed_all = n / res["latency"].max()
    log(f"[bench] batched: accuracy of {len(small)} size classes against the oracle")
    acc = _class_accuracy(W, res, paths, small, rows)
    k = fp64_sample
    log(f"[bench] batched: FP64 drop-in, one at a time, on {k} of them")
    r64, _ = _serve(pa, paths[:k], pa.PK_PREC_FP64, 1, np.zeros(k), np.arange(k), np.zeros(k, np.uint64), 0.0,
                    False)
    rmx, _ = _serve(pa, paths[:k], pa.PK_PREC_MIXED, k, np.zeros(k), np.arange(k), np.zeros(k, np.uint64),
                    0.0, False)
    fp64_rate = k / r64["latency"].max()
    mixed_rate = k / rmx["latency"].max()
    return {"workload": f"{n} instances of the batched family (sides {list(W.BATCH_SIDES)}, groups "
                        f"{list(W.BATCH_GROUPS)}, +-5% cross sections), {int(rows.sum())} rows, all "
                        f"admitted at once",
            "mixed_solves_per_s_all": mixed_all, "makespan_s_all": float(res["latency"].max()),
            "ticks_all": st["ticks"], "pool_bytes": st["pool_bytes"],
            "subset": k, "subset_rows": int(rows[:k].sum()), "fp64_dropin_solves_per_s": fp64_rate,
            "mixed_batched_solves_per_s": mixed_rate, "mixed_over_fp64": mixed_rate / fp64_rate,
            "accuracy_vs_fp64_oracle": acc, "outer_mean": float(res["outer"].mean()),
            "inner_mean": float(res["inner"].mean()), "profiled_kernels": prof, "roofline": roof,
            "spmv_all_instances": spmv, "autotuner": tuner_report(pa)}


def _sn_bases(W, tmp, n, seed=5):
    """Base S_N cases of the serving stream: the hexagonal lattice with a
    seeded number of layers (1-8) and S_N order (4, 6, 8), each with its own
    materials.  Returns (paths, rows, labels)."""
    rng = np.random.default_rng(seed)
    paths, rows, labels = [], [], []
    for j in range(n):
        layers, order = int(rng.integers(1, 9)), int(rng.choice([4, 6, 8]))
        d = os.path.join(tmp, f"sn-{j}")
        W.hex_sn_3d(d, seed=100 + j, layers=layers, order=order)
        paths.append(os.path.join(d, "input.pmp"))
        rows.append(331 * 24 * layers * 4)
        labels.append(f"S{order}, {layers} layers")
    return paths, np.array(rows), labels


def serving_numbers(pa, W, oracle, tmp, n_base=16, n_sn=4, n_req=300, rate=200.0, rel=0.02, seed=11,
                    sn_share=0.2, accuracy=True):
    """Open-loop Poisson stream of seeded +-2% perturbations of previously seen
    base cases, warm-started: 80 % diffusion requests on n_base cases of the
    batched family, 20 % S_N requests on n_sn hexagonal-lattice cases.  The
    mixed engine (continuous batching of diffusion requests, S_N requests on 4
    workers per case beside it) vs the plain FP64 engine with one request at a
    time per kind (one pool slot, one worker per S_N case), on the same
    stream."""
    paths, rows = _family(W, tmp, n_base)
    sn_paths, sn_rows, sn_labels = _sn_bases(W, tmp, n_sn)
    all_paths = paths + sn_paths
    sn_cases = set(range(n_base, n_base + n_sn))
    rng = np.random.default_rng(seed)
    arrival = np.cumsum(rng.exponential(1.0 / rate, n_req))
    arrival -= arrival[0]
    is_sn = rng.random(n_req) < sn_share
    cid = np.where(is_sn, n_base + rng.integers(0, n_sn, n_req), rng.integers(0, n_base, n_req))
    seeds = np.arange(1, n_req + 1, dtype=np.uint64)
    small = [r for r in range(n_req) if not is_sn[r] and rows[cid[r]] <= 20000][:4]
    smallest_sn = n_base + int(np.argmin(sn_rows))
    small_sn = [r for r in range(n_req) if cid[r] == smallest_sn][:2]
    out = {"workload": f"{n_req} requests, Poisson {rate:g} req/s, {int((~is_sn).sum())} diffusion on {n_base} "
                       f"base cases of the batched family ({int(rows.sum())} rows), {int(is_sn.sum())} S_N on "
                       f"{n_sn} hexagonal-lattice cases ({', '.join(sn_labels)}), +-{rel:g} seeded "
                       f"perturbations, warm start"}
    for name, prec, slots, snw in (("mixed", pa.PK_PREC_MIXED, 256, 4), ("fp64_dropin", pa.PK_PREC_FP64, 1, 1)):
        log(f"[bench] serving: {name}")
        res, st = _serve(pa, all_paths, prec, slots, arrival, cid, seeds, rel, True, keep=small + small_sn,
                         sn_workers=snw)
        span = float((arrival + res["latency"]).max())
        lat = res["latency"]
        out[name] = {"solves_per_s": n_req / span, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                     "p99_ms": float(np.percentile(lat, 99) * 1e3), "makespan_s": span,
                     "diffusion_p50_ms": float(np.percentile(lat[~is_sn], 50) * 1e3),
                     "diffusion_p99_ms": float(np.percentile(lat[~is_sn], 99) * 1e3),
                     "sn_p50_ms": float(np.percentile(lat[is_sn], 50) * 1e3) if is_sn.any() else None,
                     "sn_p99_ms": float(np.percentile(lat[is_sn], 99) * 1e3) if is_sn.any() else None,
                     "max_in_flight": st["max_live"], "converged": int(res["converged"].sum()),
                     "accuracy_vs_fp64_oracle": (_accuracy(oracle, W, res, all_paths, cid, seeds, rel,
                                                           small + small_sn, sn_cases) if accuracy else None)}
    out["mixed_over_fp64"] = out["mixed"]["solves_per_s"] / out["fp64_dropin"]["solves_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-sweep", action="store_true", help="skip the S_N sweep numbers")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-family numbers")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving numbers")
    ap.add_argument("--batch", type=int, default=256, help="instances in the batched run")
    ap.add_argument("--spmv-batched-only", action="store_true",
                    help="only the SpMV over every instance of the batched family (FP32 and FP64)")
    ap.add_argument("--serving-only", action="store_true",
                    help="only the serving stream (diffusion and S_N requests, mixed vs FP64)")
    ap.add_argument("--sweep-only", action="store_true",
                    help="only the S_N sweep timing of the hexagonal S6 lattice (FP32 and FP64)")
    ap.add_argument("--roofline-pass", action="store_true",
                    help="only the profiled solves of the headline workload (for a kernel trace whose "
                         "launches are exactly the ones the roofline figure averages)")
    ap.add_argument("--batched-pass", action="store_true",
                    help="only the profiled pass of the batched family in one mixed pool (per-kernel "
                         "rates), then the copy kernel that calibrates the traffic counters")
    ap.add_argument("--sweep-pass", action="store_true",
                    help="only the S_N sweep timing (FP32 and FP64) and the copy kernel")
    ap.add_argument("--serving-rates", type=str, default=None,
                    help="comma-separated Poisson rates (req/s): the serving stream at each rate for "
                         "--rate-window seconds of arrivals, mixed vs FP64")
    ap.add_argument("--rate-window", type=float, default=2.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import pampa_amd as pa
    from pampa_amd import workloads as W
    pa.lib()
    pa.initialize(0)
    tmp = tempfile.mkdtemp(prefix="pampa_bench_")
    d1 = W.quarter_core_2d(os.path.join(tmp, "quarter-core-170"))
    log(f"[bench] workload quarter core 170x170 G=2 at {d1}")
    if args.roofline_pass:
        print(json.dumps(roofline_pass(pa, d1)))
        return
    if args.batched_pass:
        paths, rows = _family(W, os.path.join(tmp, "family"), args.batch)
        prof, roof = batched_profile(pa, paths)
        n = 1 << 26
        sec = pa.stream_copy(n, 20)
        print(json.dumps({"profiled_kernels": prof, "roofline": roof,
                          "stream_copy": {"bytes": 8.0 * n, "us": sec * 1e6, "GBs": 8.0 * n / sec / 1e9}}))
        return
    if args.sweep_pass:
        d2 = W.hex_sn_3d(os.path.join(tmp, "hex-s6"))
        out = sweep_numbers(pa, d2)
        out["contraction_g16"] = contraction_numbers(pa, W, tmp, solves=False)
        n = 1 << 26
        sec = pa.stream_copy(n, 20)
        out["stream_copy"] = {"bytes": 8.0 * n, "us": sec * 1e6, "GBs": 8.0 * n / sec / 1e9}
        print(json.dumps(out))
        return
    if args.serving_rates:
        import oracle
        oracle.lib()
        rates = [float(r) for r in args.serving_rates.split(",")]
        out = {"window_s": args.rate_window, "rates": {}}
        for rate in rates:
            n_req = max(20, int(round(rate * args.rate_window)))
            log(f"[bench] serving at {rate:g} req/s: {n_req} requests")
            out["rates"][f"{rate:g}"] = serving_numbers(pa, W, oracle, os.path.join(tmp, f"family-{rate:g}"),
                                                        n_req=n_req, rate=rate, accuracy=False)
            r = out["rates"][f"{rate:g}"]
            log(f"[bench]   mixed {r['mixed']['solves_per_s']:.1f} solves/s p50 {r['mixed']['p50_ms']:.0f} "
                f"p99 {r['mixed']['p99_ms']:.0f} ms; fp64 {r['fp64_dropin']['solves_per_s']:.1f} solves/s")
        print(json.dumps(out))
        return
    if args.sweep_only:
        d2 = W.hex_sn_3d(os.path.join(tmp, "hex-s6"))
        out = sweep_numbers(pa, d2)
        out["eigen"] = sn_solve_numbers(pa, d2)
        print(json.dumps(out))
        return
    if args.serving_only:
        import oracle
        oracle.lib()
        print(json.dumps(serving_numbers(pa, W, oracle, os.path.join(tmp, "family"))))
        return
    if args.spmv_batched_only:
        paths, rows = _family(W, os.path.join(tmp, "family"), args.batch)
        out = batched_spmv(pa, paths)
        out["autotuner"] = tuner_report(pa)
        print(json.dumps(out))
        return

    results = {}
    for prec in ("mixed", "fp64"):
        log(f"[bench] {prec}: {args.warmup} warm-up + {args.steps} timed solves")
        results[prec] = run_solves(pa, d1, prec, args.steps, args.warmup)

    # FP64 oracle (not timed here) for the accuracy fields
    log("[bench] oracle reference + CPU baseline")
    import oracle
    oracle.lib()
    if args.no_cpu:
        cwd = os.getcwd()
        os.chdir(d1)
        ref = oracle.solve_input("input.pmp")
        os.chdir(cwd)
        cpu = None
    else:
        times, ref = cpu_baseline(d1)
        log("[bench] CPU baseline on all host cores")
        allc, procs, nsol = cpu_baseline_all_cores(d1)
        cpu = {"value": len(times) / sum(times), "unit": "solves/s", "cores": 1, "kind": "port",
               "sample": f"{len(times)} cold solves of the same 170x170 G=2 quarter core "
                         f"(assembly + sparse LU + shift-invert Arnoldi + normalisation), "
                         f"1 thread, host has {os.cpu_count()} logical CPUs",
               "all_cores": {"value": allc, "unit": "solves/s", "cores": procs, "kind": "port",
                             "sample": f"{nsol} cold solves of the same workload in {procs} processes at "
                                       f"once (1 thread each, ~10 s), aggregate rate"}}

    acc, lat = {}, {}
    for prec, (ls, k, phi, st, prof, spmv) in results.items():
        acc[prec] = {"dk_pcm": abs(k - ref.keff) * 1e5,
                     "flux_l2": float(np.linalg.norm(phi - ref.phi) / np.linalg.norm(ref.phi)),
                     "keff": k, "outer": st["outer"], "inner": st["inner"]}
        lat[prec] = {"p50_ms": float(np.percentile(ls, 50) * 1e3),
                     "p99_ms": float(np.percentile(ls, 99) * 1e3)}
    ls_mixed = results["mixed"][0]
    ls_fp64 = results["fp64"][0]
    value = len(ls_mixed) / sum(ls_mixed)
    fp64_value = len(ls_fp64) / sum(ls_fp64)

    # dominant kernel of the mixed solve, from the profiled pass
    prof = results["mixed"][4]
    dom = max(prof.items(), key=lambda kv: kv[1]["seconds"])
    dname, dv = dom
    avg_s = dv["seconds"] / dv["launches"]
    avg_b = dv["bytes"] / dv["launches"]
    achieved = avg_b / avg_s / 1e9
    roofline = {"bound": "hbm", "kernel": dname, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": SPMV_DOT_TRAFFIC_BYTES if dname == "gmres_spmv_dot" else None,
                "avg_launch_us": avg_s * 1e6, "bytes_per_launch": avg_b,
                "launches_per_solve": dv["launches"],
                "note": "latency-bound: the 170x170 operator (4 MB) is cache resident; every launch of "
                        "the captured cycles counted, idle ones included (HIP events recorded by each "
                        "launch, same kernels as the graph)"}
    spmv_s, spmv_b = results["mixed"][5]
    spmv = {"kernel": "spmv_fp32", "us": spmv_s * 1e6, "GBs": spmv_b / spmv_s / 1e9,
            "frac_hbm": spmv_b / spmv_s / 1e9 / HBM_PEAK_GBS, "bytes": spmv_b}

    sweep = None
    if not args.no_sweep:
        d2 = W.hex_sn_3d(os.path.join(tmp, "hex-s6"))
        log("[bench] S6 hex lattice sweep timing")
        sweep = sweep_numbers(pa, d2)
        sweep["workload"] = "hex lattice 331 columns x 24 triangles x 40 layers, G=4, S6"
        log("[bench] S6 hex lattice eigen solves")
        sweep["eigen"] = sn_solve_numbers(pa, d2)
        log("[bench] 16-group source contraction, vector kernel vs matrix cores")
        sweep["contraction_g16"] = contraction_numbers(pa, W, tmp)

    batched = serving = None
    if not args.no_batched:
        batched = batched_numbers(pa, W, oracle, os.path.join(tmp, "family"), n_inst=args.batch)
    if not args.no_serving:
        serving = serving_numbers(pa, W, oracle, os.path.join(tmp, "family"))

    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 / value, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 inner / f64 outer",
        "data": "synthetic (seeded SplitMix64 .pmp inputs)",
        "config": {"workload": "quarter-core 170x170 cells (1 cm), 2 groups, 4 materials, "
                               "k-eigen residual 1e-8, cold start per solve",
                   "unknowns": 57800, "nnz_R": 345440},
        "fp64_value": fp64_value, "mixed_over_fp64": value / fp64_value,
        "latency": lat, "accuracy_vs_fp64_oracle": acc,
        "roofline": roofline, "spmv": spmv, "sweep": sweep, "batched": batched, "serving": serving,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

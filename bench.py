"""Benchmark of the sparse hot path on one MI355X.

Default workload (the configuration the metric is quoted on): 3-D 7-point
Laplacian on a 128^3 grid (n = 2,097,152, nnz = 14,581,760), FP64 CSR with
int32 indices, RCM-reordered on the host.

  * value: FP64 CSR SpMV y = A x throughput in GB/s of algorithmic bytes
    B = nnz*(8+4) + (n+1)*4 + n*8 (x, counted once) + n*8 (y), K timed
    launches of the kernel the default (AUTO) request runs -- the sliced-ELL
    kernel for this matrix; the padding its slices carry is read but not
    counted in B. The matrix is held in three copies used in rotation, so a
    launch never finds its matrix in the 256 MB Infinity Cache: each timed
    launch streams A from HBM.
  * pcg: IC(0)-PCG on b ~ U[-1,1) (seeded) to a 1e-10 true relative residual,
    plain FP64 (exact triangular solves, true residual every iteration: the
    reference's algorithm) vs mixed precision (FP32 or BF16 matrix and
    factors, FP32 vectors, FP64 dot products, FP64 x and FP64 true residual
    through reliable updates), each with exact (sync-free) triangular
    solves and with 3 Jacobi sweeps per triangle.
  * roofline: the SpMV kernel against the 8 TB/s HBM peak, its average
    launch duration from one pair of HIP events around the K timed launches
    on the stream they run on (so gaps between launches count too); beside
    it a measured ceiling, a 1 GiB device-to-device copy.
  * cpu_baseline: the oracle's OpenMP SpMV (a restatement of the reference's
    ParallelSPMV) on this host's cores, the median of single SpMVs timed
    over a bounded ~10 s sample after one warm-up, with the CPU model and
    core count, also in the reference's 16 B/nnz convention.
  * c4, c3, c5: the other configurations' records as extra keys of the same
    line (27-point 160^3 ILU(1)-GMRES(30) FP64 vs mixed precision with FP64
    refinement; power-law SpMV in FP64/FP32/BF16; the adjoint of a 512^2
    diffusion solve).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c1..c5]
       [--no-pcg] [--no-cpu] [--no-extra] [--pmc-probe] [--trsv level|syncfree]
       [--c4-grid G]

--config c1 (the 256^2 Poisson chain next to the CPU oracle), c3, c4 and c5
print only that configuration's record; the default (c2) prints the bench
line with the c4, c3 and c5 records (leave them out with --no-extra).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SpMV GB/s (%HBM peak); PCG time to 1e-10: mixed FP32/BF16 vs FP64, "
          "1 GPU")
HBM_PEAK_GBS = 8000.0
# SPMV_* codes (ROW 0, SEGSUM 1, ALBUS 2, AUTO 3, SELL 4)
ALG_NAMES = {0: "row", 1: "segsum", 2: "albus", 3: "auto", 4: "sell", 5: "colblock"}
ALG_DESC = {0: "one wavefront per row", 2: "nnz-balanced segment kernel, 512 nnz per "
            "wavefront, + boundary-row fix-up",
            4: "sliced ELL, 64-row column-major slices, one lane per row, 16-bit column "
               "offsets where they fit"}
ALG_KERNEL = {0: "k_spmv_row<G,RB,double,double,double>",
              2: "k_spmv_seg<double,double,double> + k_spmv_seg_fix (one SpMV; duration "
                 "of the pair)",
              4: "k_spmv_sell<%s,double,double,double>"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, s_v=8, s_x=8, s_y=8, beta_nonzero=False):
    return nnz * (s_v + 4) + (n + 1) * 4 + n * s_x + n * s_y * (2 if beta_nonzero else 1)


def mixed_summary(sec, groups):
    """Mixed precision against FP64 with the same preconditioner (same
    triangular-solve mode and sweep count): per group, the FP64 time, the
    fastest mixed variant and the ratio. `groups` maps a label to (the FP64
    variant, the mixed variants)."""
    out = {}
    for label, (ref, mixed) in groups.items():
        mixed = [k for k in mixed if k in sec]
        if ref not in sec or not mixed:
            continue
        best = min(mixed, key=lambda k: sec[k])
        out[label] = {"fp64_s": sec[ref], "best_mixed": best, "mixed_s": sec[best],
                      "speedup": round(sec[ref] / sec[best], 3),
                      "all": {k: round(sec[ref] / sec[k], 3) for k in mixed}}
    return out


def build_c2(la, grid=128, rcm=True):
    t0 = time.time()
    A = la.gen.laplace3d(grid)
    t_gen = time.time() - t0
    t_rcm = t_perm = 0.0
    if rcm:
        t0 = time.time()
        ip, p = la.rcm(A)
        t_rcm = time.time() - t0
        t0 = time.time()
        A = la.permute(A, ip, p)
        t_perm = time.time() - t0
    return A, {"gen_s": round(t_gen, 3), "rcm_s": round(t_rcm, 3), "permute_s": round(t_perm, 3)}


def time_spmv(la, torch, mats, x, y, steps, warmup, alg, val, acc, stream):
    """Average launch duration (ms) of K back-to-back launches, from one pair
    of HIP events around them on the stream they run on (launch gaps
    included), and the host wall time of the K launches (ms); rotating over
    `mats`."""
    for i in range(warmup):
        mats[i % len(mats)].spmv(x, y, alg=alg, val=val, acc=acc)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(steps):
        mats[i % len(mats)].spmv(x, y, alg=alg, val=val, acc=acc)
    e1.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    return [e0.elapsed_time(e1) / steps], wall_ms


def copy_ceiling(torch, dev, mib=1024, reps=10):
    """Measured bandwidth ceilings next to the 8 TB/s peak: a device-to-device
    copy of `mib` MiB (that many bytes read and as many written) and a sum
    over the same buffer (read only, as an SpMV nearly is), each timed one
    call at a time with HIP events; the medians."""
    src = torch.empty(mib << 17, dtype=torch.float64, device=dev)
    src.fill_(1.0)
    dst = torch.empty_like(src)

    def med(fn):
        fn()
        torch.cuda.synchronize()
        times = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1) * 1e-3)
        return float(np.median(times))

    t = med(lambda: dst.copy_(src))
    tr = med(lambda: torch.sum(src))
    nbytes, rbytes = 2 * (mib << 20), mib << 20
    del src, dst
    return {"what": "%d MiB device-to-device copy (read + write), median of %d" % (mib, reps),
            "bytes": nbytes, "us": round(t * 1e6, 2), "GBs": round(nbytes / t / 1e9, 1),
            "frac_of_peak": round(nbytes / t / 1e9 / HBM_PEAK_GBS, 4),
            "read": {"what": "torch.sum over the same %d MiB (read only)" % mib,
                     "us": round(tr * 1e6, 2), "GBs": round(rbytes / tr / 1e9, 1),
                     "frac_of_peak": round(rbytes / tr / 1e9 / HBM_PEAK_GBS, 4)}}


def cpu_info():
    """Model string, logical CPUs this process may run on, and the physical
    cores among them (distinct (package, core) pairs in /proc/cpuinfo)."""
    if _CPU_INFO:
        return _CPU_INFO
    model, cores, cur, ident = "unknown", set(), {}, {}
    allowed = os.sched_getaffinity(0)
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                cur = {"cpu": int(v)}
            elif k == "model name":
                model = v
            elif k in ("cpu family", "model", "stepping"):
                ident[k.replace(" ", "_")] = v
            elif k == "physical id":
                cur["pkg"] = v
            elif k == "core id" and cur.get("cpu") in allowed:
                cores.add((cur.get("pkg"), v))
    except OSError:
        pass
    _CPU_INFO.update({"model": model, "ident": ident, "logical_cpus": len(allowed),
                      "physical_cores": len(cores) or None})
    return _CPU_INFO


# read at start-up: an OpenMP runtime started with OMP_PROC_BIND set binds the
# main thread to one core, which would hide the process's CPU set
_CPU_INFO = {}


def cpu_child(path, seconds, min_runs=10):
    """The CPU baseline's timing, in a process of its own started with
    OMP_PROC_BIND=close and OMP_PLACES=cores: the oracle's OpenMP
    ParallelSPMV restatement on the matrix saved at `path`; one warm-up,
    then single SpMVs timed one by one until `seconds` have passed (at
    least `min_runs`). Prints the times as one JSON line."""
    from oracle import oracle as orc
    d = np.load(path)
    Ao = orc.CSR(int(d["n"]), d["ai"], d["aj"], d["av"])
    x = np.random.default_rng(1).uniform(-1, 1, Ao.n)
    threads = orc.max_threads()
    orc.spmv(Ao, x, alg="parallel", nthreads=threads)  # warm-up
    times, t_start = [], time.perf_counter()
    while len(times) < min_runs or time.perf_counter() - t_start < seconds:
        t0 = time.perf_counter()
        orc.spmv(Ao, x, alg="parallel", nthreads=threads)
        times.append(time.perf_counter() - t0)
        if len(times) >= 100000:
            break
    print(json.dumps({"times": times, "threads": threads}))


def cpu_baseline(A, seconds=10.0):
    """The oracle's OpenMP ParallelSPMV restatement on this host's cores,
    timed in a child process with its threads packed onto cores
    (OMP_PROC_BIND=close, OMP_PLACES=cores; this process's own threads stay
    unbound); the median run is reported."""
    import subprocess
    import tempfile
    ci = cpu_info()
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        path = os.path.join(tmp, "a.npz")
        np.savez(path, n=np.int64(A.n), ai=np.asarray(A.ai), aj=np.asarray(A.aj),
                 av=np.asarray(A.av))
        # the thread count the environment sets (the host's CPU share), else
        # one per logical CPU this process may use
        env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
        env.setdefault("OMP_NUM_THREADS", str(ci["logical_cpus"]))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", path,
                            "--cpu-seconds", str(seconds)], env=env, capture_output=True,
                           text=True, check=True)
    out = json.loads(r.stdout.strip().splitlines()[-1])
    times, threads = out["times"], out["threads"]
    med = float(np.median(times))
    B = spmv_bytes(A.n, A.nnz)
    return {"value": round(B / med / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "GBs_16B_per_nnz": round(16.0 * A.nnz / med / 1e9, 3),
            "median_ms": round(med * 1e3, 4),
            "p10_p90_ms": [round(float(np.percentile(times, q)) * 1e3, 4) for q in (10, 90)],
            "runs": len(times),
            "sample": "%d single FP64 SpMVs after one warm-up (oracle ParallelSPMV "
                      "restatement, OpenMP, %d threads, OMP_PROC_BIND=close "
                      "OMP_PLACES=cores) on the same RCM-ordered 128^3 7-point matrix, "
                      "%.1f s; value = median" % (len(times), threads, sum(times)),
            "cpu_model": ci["model"], "cpu_ident": ci["ident"],
            "logical_cpus": ci["logical_cpus"], "physical_cores": ci["physical_cores"]}


def run_pcg(la, torch, A, dev, stream, trsv_mode, variants, cpu=False):
    U = la.split_triangle(A, upper=True)
    t0 = time.time()
    R, ok = la.icc(U, 0)
    t_fact = time.time() - t0
    if not ok:
        raise RuntimeError("IC(0) failed")
    t0 = time.time()
    P = la.Preconditioner(R, "ic")
    P.set_stream(stream)
    t_levels = time.time() - t0
    dA = la.DeviceMatrix(A, stream=stream)
    b = torch.tensor(la.gen.uniform(A.n, 2024), device=dev)
    # the same IC(0) with the numeric phase on the device (bitwise the host's)
    t0 = time.time()
    Pg = la.Preconditioner.ic_gpu(A, 0)
    t_gsetup = time.time() - t0
    av = torch.tensor(A.av, device=dev)
    Pg.refactor(av)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Pg.refactor(av)
    torch.cuda.synchronize()
    t_gnum = time.perf_counter() - t0
    same = bool(np.array_equal(Pg.factor().av, R.av))
    del Pg
    P.set_sweeps(3)
    out = {"preconditioner": "IC(0)", "levels": P.levels(), "factor_s": round(t_fact, 3),
           "analysis_s": round(t_levels, 3),
           "device_factor": {"setup_s": round(t_gsetup, 3), "numeric_s": round(t_gnum, 5),
                             "bitwise_equal_host": same},
           "trsv_mode": ["level", "syncfree"][trsv_mode], "runs": {}}
    xs_all = {}
    for name, kw in variants:
        kw = dict(kw)
        kw.setdefault("trsv_mode", trsv_mode)
        P.set_sweeps(kw.pop("sweeps", 3))
        x = torch.zeros(A.n, dtype=torch.float64, device=dev)
        # one untimed run per variant so level graphs and precision copies
        # exist; then the timed one
        la.pcg(dA, P, b, x, rel_tol=1e-10, max_iters=5000, **kw)
        x.zero_()
        torch.cuda.synchronize()
        res = la.pcg(dA, P, b, x, rel_tol=1e-10, max_iters=5000, **kw)
        xs_all[name] = x.cpu().numpy()
        if name == "fp64":
            x_fp64 = xs_all[name]
        out["runs"][name] = {"seconds": round(res["seconds"], 5), "iters": res["iters"],
                             "outer": res["outer_iters"],
                             "rel_residual": float("%.3e" % res["rel_residual"]),
                             "converged": res["converged"]}
        log("pcg", name, out["runs"][name])
    if cpu:
        # the same solve by the oracle's serial FP64 PCG (the reference's
        # solver loop restated, one core), for comparison
        from oracle import oracle as orc
        Ao = orc.CSR(A.n, A.ai, A.aj, A.av)
        Ro = orc.CSR(R.n, R.ai, R.aj, R.av)
        bh = b.cpu().numpy()
        t0 = time.perf_counter()
        xo, conv, it, rr = orc.pcg(Ao, bh, kind="ic", F=Ro, rel_tol=1e-10, maxiter=5000)
        el = time.perf_counter() - t0
        out["cpu_fp64"] = {"seconds": round(el, 3), "iters": it, "rel_residual": float("%.3e" % rr),
                           "converged": bool(conv), "cores": 1,
                           "kind": "port (oracle PCG + IC(0) apply, serial)",
                           "gpu_fp64_x_rel_diff": float("%.3e" % (
                               np.linalg.norm(x_fp64 - xo) / np.linalg.norm(xo)))}
        # every GPU solve's solution against the CPU oracle's FP64 one (all
        # stop at a 1e-10 relative residual; cond(A) ~ 6.6e3 bounds the gap)
        for k, xk in xs_all.items():
            out["runs"][k]["x_rel_diff_vs_cpu_oracle"] = float(
                "%.3e" % (np.linalg.norm(xk - xo) / np.linalg.norm(xo)))
        log("pcg cpu", out["cpu_fp64"])
    return out


def _spmv_table(la, torch, A, mats, dev, stream, algs, dtypes, steps, seed=77):
    """GB/s and relative error against the FP64 oracle for each (alg, dtype).
    GBs counts the CSR algorithmic bytes for that value type; where AUTO
    runs the sliced-ELL copy, layout_GBs / layout_frac count the bytes that
    copy streams itself (every slot's value and column offset, the slice
    table, x once and y)."""
    from oracle import oracle as orc
    n, nnz = A.n, A.nnz
    sinfo = mats[0].sell_info()
    sell = mats[0].resolve(la.SPMV_AUTO) == 4 and sinfo["built"]
    x64 = torch.tensor(la.gen.uniform(n, seed), device=dev)
    ref = orc.spmv(orc.CSR(n, A.ai, A.aj, A.av), x64.cpu().numpy(), alg="parallel",
                   nthreads=orc.max_threads())
    out = {}
    for dt, name in dtypes:
        for m in mats:
            if dt != la.F64:
                m.add_dtype(dt)
        vec = torch.float64 if dt == la.F64 else torch.float32
        acc = la.F64 if dt == la.F64 else la.F32
        x = x64.to(vec)
        y = torch.zeros(n, dtype=vec, device=dev)
        for aname, alg in algs:
            mats[0].spmv(x, y, alg=alg, val=dt, acc=acc)
            torch.cuda.synchronize()
            err = float(np.linalg.norm(y.double().cpu().numpy() - ref) / np.linalg.norm(ref))
            d, _ = time_spmv(la, torch, mats, x, y, steps, 3, alg, dt, acc, stream)
            sv = {la.F64: 8, la.F32: 4, la.BF16: 2}[dt]
            sx = 8 if dt == la.F64 else 4
            B = spmv_bytes(n, nnz, s_v=sv, s_x=sx, s_y=sx)
            us = float(np.mean(d)) * 1e3
            rec = {"us": round(us, 2), "GBs": round(B / us / 1e3, 1),
                   "rel_err_vs_fp64": float("%.3e" % err)}
            if alg == la.SPMV_AUTO and sell:
                Bl = (sinfo["slots"] * (sv + (2 if sinfo["col16"] else 4)) + (n + 63) // 64 * 12
                      + 2 * n * sx)
                rec["layout_GBs"] = round(Bl / us / 1e3, 1)
                rec["layout_frac"] = round(Bl / us / 1e3 / HBM_PEAK_GBS, 4)
            out["%s_%s" % (aname, name)] = rec
            log(aname, name, out["%s_%s" % (aname, name)])
    return out


def bench_c1(la, torch, dev, stream, args):
    """The reference's CPU-runnable chain: 2-D 5-point Poisson 256^2
    (n = 65,536), SpMV y = A x (x ~ U[-1,1), seeded), ILU(0) factorisation
    and its forward/backward solve, IC(0)-PCG on b = A 1 to a 1e-8 true
    relative residual; each step on the GPU (device-resident, FP64) next to
    the oracle's serial restatement on one host core, with the GPU result's
    relative difference from the oracle's."""
    from oracle import oracle as orc
    A = la.gen.poisson2d(256)
    Ao = orc.CSR(A.n, A.ai, A.aj, A.av)
    xh = la.gen.uniform(A.n, 101)
    bh = orc.spmv(Ao, np.ones(A.n))
    out = {"config": "c1 poisson2d 256^2", "n": A.n, "nnz": A.nnz}

    def gpu_time(fn, reps=20):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    def cpu_time(fn, reps=5):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        return (time.perf_counter() - t0) / reps, r
    d = la.DeviceMatrix(A, stream=stream)
    x, y = torch.tensor(xh, device=dev), torch.zeros(A.n, dtype=torch.float64, device=dev)
    t_gpu = gpu_time(lambda: d.spmv(x, y))
    t_cpu, yr = cpu_time(lambda: orc.spmv(Ao, xh))
    out["spmv"] = {"gpu_us": round(t_gpu * 1e6, 2), "cpu_us": round(t_cpu * 1e6, 1),
                   "rel_diff": float("%.2e" % (np.linalg.norm(y.cpu().numpy() - yr) /
                                               np.linalg.norm(yr)))}
    t0 = time.perf_counter()
    F = la.iluk(A, 0)
    t_fact = time.perf_counter() - t0
    P = la.Preconditioner(F, "ilu")
    P.set_stream(stream)
    b = torch.tensor(bh, device=dev)
    z = torch.zeros_like(b)
    Fo = orc.CSR(F.n, F.ai, F.aj, F.av)
    t_gpu = gpu_time(lambda: P.apply(b, z))
    t_cpu, zr = cpu_time(lambda: orc.prec_apply(Ao, "ilu", Fo, bh))
    out["ilu0"] = {"host_factor_s": round(t_fact, 4), "levels_L_U": P.levels(),
                   "apply_gpu_us": round(t_gpu * 1e6, 1), "apply_cpu_us": round(t_cpu * 1e6, 1),
                   "rel_diff": float("%.2e" % (np.linalg.norm(z.cpu().numpy() - zr) /
                                               np.linalg.norm(zr)))}
    R, ok = la.icc(la.split_triangle(A, upper=True), 0)
    Pc = la.Preconditioner(R, "ic")
    Pc.set_stream(stream)
    Ro = orc.CSR(R.n, R.ai, R.aj, R.av)
    runs = {}
    for name, kw in [("fp64", {}), ("mixed_bf16_jacobi3",
                                    dict(precision=la.PREC_MIXED_BF16,
                                         trsv_mode=la.TRSV_JACOBI))]:
        xs = torch.zeros_like(b)
        la.pcg(d, Pc, b, xs, rel_tol=1e-8, max_iters=2000, **kw)
        xs.zero_()
        torch.cuda.synchronize()
        r = la.pcg(d, Pc, b, xs, rel_tol=1e-8, max_iters=2000, **kw)
        runs[name] = {"seconds": round(r["seconds"], 5), "iters": r["iters"],
                      "rel_residual": float("%.2e" % r["rel_residual"]),
                      "x_minus_ones": float("%.2e" % (np.linalg.norm(xs.cpu().numpy() - 1.0) /
                                                       np.sqrt(A.n)))}
    t0 = time.perf_counter()
    xo, conv, it, rr = orc.pcg(Ao, bh, kind="ic", F=Ro, rel_tol=1e-8, maxiter=2000)
    runs["cpu_oracle_fp64_serial"] = {"seconds": round(time.perf_counter() - t0, 4), "iters": it,
                                      "rel_residual": float("%.2e" % rr)}
    out["ic0_pcg"] = runs
    return out


def bench_c3(la, torch, dev, stream, args):
    """Power-law matrix, n = 4M, lognormal row lengths (mean 16, sigma 1,
    clipped to 20000), uniform deduplicated columns: SpMV load balance."""
    t0 = time.time()
    A = la.gen.powerlaw(4_000_000, 16.0, 1.0, 20000, seed=3)
    prep = round(time.time() - t0, 3)
    if args.pmc_probe:  # 4 FP64 SpMVs of each form + a 1 GiB copy, for rocprofv3 --pmc
        d = la.DeviceMatrix(A, stream=stream)
        x = torch.tensor(la.gen.uniform(A.n, 7), device=dev)
        y = torch.zeros_like(x)
        for alg in (la.SPMV_ALBUS, la.SPMV_AUTO):
            for _ in range(4):
                d.spmv(x, y, alg=alg)
        big = torch.empty(1 << 27, dtype=torch.float64, device=dev)
        torch.empty_like(big).copy_(big)
        torch.cuda.synchronize()
        return {"probe": "c3 spmv albus x4, auto (row panels x column blocks) x4, copy 1GiB", "nnz": A.nnz}
    rl = np.diff(A.ai)
    mats = [la.DeviceMatrix(A, stream=stream) for _ in range(2)]
    algs = [("auto", la.SPMV_AUTO), ("albus", la.SPMV_ALBUS), ("row", la.SPMV_ROW)]
    # AUTO picks the row-panel x column-block kernel here: one launch, each
    # workgroup sweeps blocks of 65,536 columns with its rows' sums in LDS
    tab = _spmv_table(la, torch, A, mats, dev, stream, algs,
                      [(la.F64, "f64"), (la.F32, "f32"), (la.BF16, "bf16")], 20)
    return {"config": "c3 power-law n=4M", "n": A.n, "nnz": A.nnz,
            "auto_runs": ALG_NAMES[mats[0].resolve(la.SPMV_AUTO)],
            "max_row": int(rl.max()), "gen_s": prep, "spmv": tab,
            "note": "f32/bf16: values in that type, FP32 vectors and accumulation; GBs of "
                    "the CSR algorithmic bytes for that value type"}


def bench_c4(la, torch, dev, stream, args):
    """27-point anisotropic variable-coefficient diffusion, 160^3: SpMV in
    FP64/FP32/BF16, then ILU(1)-GMRES(30) to 1e-10, plain FP64 against
    mixed precision with FP64 iterative refinement."""
    grid = args.c4_grid
    t0 = time.time()
    A = la.gen.aniso27(grid)
    t_gen = time.time() - t0
    mats = [la.DeviceMatrix(A, stream=stream)]
    tab = _spmv_table(la, torch, A, mats, dev, stream,
                      [("auto", la.SPMV_AUTO), ("albus", la.SPMV_ALBUS)],
                      [(la.F64, "f64"), (la.F32, "f32"), (la.BF16, "bf16")], 10)
    t0 = time.time()
    fai, faj, fu = la.iluk_symbolic(A, 1)
    t_sym = time.time() - t0
    t0 = time.time()
    F = la.iluk_numeric(A, fai, faj, fu)
    t_num = time.time() - t0
    t0 = time.time()
    P = la.Preconditioner(F, "ilu")
    P.set_stream(stream)
    t_an = time.time() - t0
    # ILU(1) with the numeric phase on the device
    t0 = time.time()
    Pg = la.Preconditioner.ilu_gpu(A, 1)
    t_gsetup = time.time() - t0
    av = torch.tensor(A.av, device=dev)
    Pg.refactor(av)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Pg.refactor(av)
    torch.cuda.synchronize()
    t_gnum = time.perf_counter() - t0
    same = bool(np.array_equal(Pg.factor().av, F.av))
    del Pg, av
    b = torch.tensor(la.gen.uniform(A.n, 4), device=dev)
    mode = la.TRSV_LEVEL if args.trsv == "level" else la.TRSV_SYNCFREE
    runs = {}
    # the configuration's two mixed forms: FP32 factors with the matrix in
    # FP32 or BF16; plus everything in BF16
    # and with 6 Jacobi sweeps per triangle in place of the exact solves
    P.set_sweeps(6)
    J = dict(trsv_mode=la.TRSV_JACOBI)
    variants = [("fp64", dict(precision=la.PREC_FP64)),
                ("mixed_f32", dict(precision=la.PREC_MIXED_F32, inner_tol=1e-5)),
                ("mixed_bf16_matrix_f32_factors", dict(precision=la.PREC_MIXED_BF16_F32P,
                                                       inner_tol=1e-3)),
                ("mixed_bf16", dict(precision=la.PREC_MIXED_BF16, inner_tol=1e-3)),
                ("fp64_jacobi6", dict(precision=la.PREC_FP64, **J)),
                ("mixed_f32_jacobi6", dict(precision=la.PREC_MIXED_F32, inner_tol=1e-4, **J)),
                ("mixed_bf16_matrix_f32_factors_jacobi6",
                 dict(precision=la.PREC_MIXED_BF16_F32P, inner_tol=1e-3, **J)),
                # beyond the configuration's FP32 factors: BF16 factors with
                # the matrix kept in FP32
                ("mixed_f32_matrix_bf16_factors_jacobi6",
                 dict(precision=la.PREC_MIXED_F32_BF16P, inner_tol=1e-4, **J))]
    for name, kw in variants:
        kw = dict(kw)
        kw.setdefault("trsv_mode", mode)
        x = torch.zeros(A.n, dtype=torch.float64, device=dev)
        la.fgmres(mats[0], P, b, x, rel_tol=1e-10, restart=30, max_iters=3000,
                  **kw)  # untimed: builds precision copies
        x.zero_()
        torch.cuda.synchronize()
        r = la.fgmres(mats[0], P, b, x, rel_tol=1e-10, restart=30, max_iters=3000, **kw)
        runs[name] = {"seconds": round(r["seconds"], 4), "iters": r["iters"],
                      "outer": r["outer_iters"], "rel_residual": float("%.3e" % r["rel_residual"]),
                      "converged": r["converged"]}
        xh = x.cpu().numpy()
        if name == "fp64":
            x64 = xh
        # against the plain FP64 GPU solve (a serial CPU FGMRES with this
        # ILU(1) at 4.1M unknowns would take minutes)
        runs[name]["x_rel_diff_vs_gpu_fp64"] = float(
            "%.3e" % (np.linalg.norm(xh - x64) / np.linalg.norm(x64)))
        log("gmres", name, runs[name])
    ap = {}
    r = torch.tensor(la.gen.uniform(A.n, 5), device=dev)
    z = torch.zeros_like(r)
    for val, vn in [(la.F64, "f64"), (la.F32, "f32")]:
        for md, mn in [(mode, args.trsv), (la.TRSV_JACOBI, "jacobi6")]:
            rr = r if val == la.F64 else r.float()
            zz = z if val == la.F64 else z.float()
            P.apply(rr, zz, mode=md, val=val)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                P.apply(rr, zz, mode=md, val=val)
            torch.cuda.synchronize()
            ap["%s_%s" % (mn, vn)] = round((time.perf_counter() - t0) / 5 * 1e3, 3)
    out = {"config": "c4 aniso27 %d^3" % grid, "n": A.n, "nnz": A.nnz, "spmv": tab,
           "auto_runs": ALG_NAMES[mats[0].resolve(la.SPMV_AUTO)],
           "ilu1": {"nnz": int(len(faj)), "levels_L_U": P.levels(), "gen_s": round(t_gen, 2),
                    "symbolic_s": round(t_sym, 2), "numeric_s": round(t_num, 2),
                    "analysis_upload_s": round(t_an, 2), "apply_ms": ap,
                    "device_numeric": {"setup_s": round(t_gsetup, 2),
                                       "numeric_s": round(t_gnum, 4),
                                       "bitwise_equal_host": same}},
           "trsv_mode": args.trsv, "gmres30": runs}
    sec = {k: v["seconds"] for k, v in runs.items()}
    out["mixed_vs_fp64"] = mixed_summary(
        sec, {"exact (%s)" % args.trsv: ("fp64", [k for k in sec if k.startswith("mixed")
                                                   and "jacobi" not in k]),
              "6 Jacobi sweeps": ("fp64_jacobi6", [k for k in sec if k.startswith("mixed")
                                                   and "jacobi" in k])})
    return out


def bench_c5(la, torch, dev, stream, args):
    """Adjoint of x = A(theta)^-1 b, 2-D 5-point edge-conductivity diffusion
    on a 512^2 grid, theta ~ lognormal(0, 0.5) (seeded): loss ||x - x*||^2
    with x* = A(theta*)^-1 b; one forward IC(0)-PCG solve, the transposed
    (adjoint) solve and the value gradient, timed per variant, with each
    gradient's relative error against an FP64 CPU reference (sparse direct
    LU and the closed form dL/dtheta_e = -(lambda_i - lambda_j)(x_i - x_j))."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    from linearalgebra_amd.autograd import SparseSystem, edge_values, sparse_solve
    nx = 512
    ne = 2 * nx * nx - 2 * nx
    theta = la.gen.lognormal(ne, 5, 0.0, 0.5)
    theta_star = la.gen.lognormal(ne, 55, 0.0, 0.5)
    A0, ei, ej, vm = la.gen.diffusion2d(nx, nx, np.zeros(ne))
    bh = la.gen.uniform(nx * nx, 505)
    vm_t = torch.tensor(vm, dtype=torch.int64, device=dev)
    base = torch.tensor(A0.av, device=dev)
    # FP64 CPU reference gradient
    t0 = time.time()
    A, _, _, _ = la.gen.diffusion2d(nx, nx, theta)
    As, _, _, _ = la.gen.diffusion2d(nx, nx, theta_star)

    def lu(M):
        return spl.splu(sp.csr_matrix((M.av, M.aj, M.ai), shape=(M.n, M.n)).tocsc(),
                        permc_spec="MMD_AT_PLUS_A")
    LA = lu(A)
    x_ref = LA.solve(bh)
    lam_ref = LA.solve(2 * (x_ref - lu(As).solve(bh)))
    g_ref = -(lam_ref[ei] - lam_ref[ej]) * (x_ref[ei] - x_ref[ej])
    t_ref = time.time() - t0
    J = dict(trsv_mode=la.TRSV_JACOBI)
    variants = [("fp64", dict(precision=la.PREC_FP64)),
                ("mixed_f32", dict(precision=la.PREC_MIXED_F32)),
                ("fp64_jacobi3", dict(precision=la.PREC_FP64, **J)),
                ("mixed_f32_jacobi3", dict(precision=la.PREC_MIXED_F32, **J)),
                ("mixed_bf16_jacobi3", dict(precision=la.PREC_MIXED_BF16, **J))]
    runs = {}
    for name, kw in variants:
        system = SparseSystem(A0, kind="ic", level=0, rel_tol=1e-11, max_iters=20000,
                              adjoint_rel_tol=1e-10, **kw)
        system.P.set_sweeps(3)

        def once():
            b = torch.tensor(bh, device=dev, requires_grad=True)
            th = torch.tensor(theta, device=dev, requires_grad=True)
            with torch.no_grad():
                xs = sparse_solve(edge_values(torch.tensor(theta_star, device=dev), base, vm_t),
                                  b.detach(), system)
            x = sparse_solve(edge_values(th, base, vm_t), b, system)
            loss = ((x - xs) ** 2).sum()
            loss.backward()
            return th.grad, b.grad
        try:
            once()  # untimed: precision copies, Jacobi forms, transposed plans
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            g, gb = once()
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
        except la.LaError as e:  # reported, not hidden: a precision too coarse
            runs[name] = {"converged": False, "error": str(e)}
            log("c5", name, runs[name])
            continue
        g, gb = g.cpu().numpy(), gb.cpu().numpy()
        runs[name] = {
            "seconds_two_forward_one_adjoint": round(el, 4),
            "forward_iters": system.last["forward"]["iters"],
            "adjoint_iters": system.last["adjoint"]["iters"],
            "grad_theta_rel_err": float("%.3e" % (np.linalg.norm(g - g_ref) /
                                                 np.linalg.norm(g_ref))),
            "grad_b_rel_err": float("%.3e" % (np.linalg.norm(gb - lam_ref) /
                                             np.linalg.norm(lam_ref))),
            "converged": True}
        log("c5", name, runs[name])
    sec = {k: v["seconds_two_forward_one_adjoint"] for k, v in runs.items()
           if "seconds_two_forward_one_adjoint" in v}
    return {"config": "c5 adjoint 2-D diffusion 512^2", "n": nx * nx,
            "edges": ne, "cpu_reference_s": round(t_ref, 2), "runs": runs,
            "mixed_vs_fp64": mixed_summary(
                sec, {"exact (syncfree)": ("fp64", [k for k in sec if k.startswith("mixed")
                                                     and "jacobi" not in k]),
                      "3 Jacobi sweeps": ("fp64_jacobi3", [k for k in sec if k.startswith("mixed")
                                                           and "jacobi" in k])})}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-pcg", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only a few SpMV launches + a 1 GiB copy (for rocprofv3 --pmc)")
    ap.add_argument("--trsv", default="syncfree", choices=["level", "syncfree"])
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--c4-grid", type=int, default=160)
    ap.add_argument("--no-extra", action="store_true",
                    help="leave out the c4, c3 and c5 records of the default line")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:  # the CPU baseline's own process (no GPU)
        cpu_child(args.cpu_child, args.cpu_seconds)
        return
    cpu_info()  # before torch or the libraries start an OpenMP runtime
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; the multi-GPU sharding of the "
              "path is a design note (row-block halo exchange), not a measured mode.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import linearalgebra_amd as la
    la.lib()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    others = {"c1": bench_c1, "c3": bench_c3, "c4": bench_c4, "c5": bench_c5}
    if args.config in others:
        print(json.dumps(others[args.config](la, torch, dev, stream, args)))
        return
    if args.config != "c2":
        raise SystemExit("--config is one of c2 (the bench line), c1, c3, c4, c5")
    A, prep = build_c2(la)
    n, nnz = A.n, A.nnz
    log("matrix n=%d nnz=%d prep=%s" % (n, nnz, prep))
    mats = [la.DeviceMatrix(A, stream=stream) for _ in range(3)]
    x = torch.tensor(la.gen.uniform(n, 77), device=dev)
    y = torch.zeros(n, dtype=torch.float64, device=dev)

    if args.pmc_probe:
        for alg in (la.SPMV_AUTO, la.SPMV_ALBUS, la.SPMV_ROW):
            for i in range(6):
                mats[i % 3].spmv(x, y, alg=alg)
        big = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        dst = torch.empty_like(big)
        dst.copy_(big)
        torch.cuda.synchronize()
        print(json.dumps({"probe": "spmv auto (sliced ELL) x6, albus x6, row x6 + copy 1GiB",
                          "spmv_bytes": spmv_bytes(n, nnz)}))
        return

    # parity check of the timed kernel against the FP64 oracle (cheap)
    from oracle import oracle as orc
    chosen = mats[0].resolve(la.SPMV_AUTO)
    sinfo = mats[0].sell_info()
    mats[0].spmv(x, y, alg=la.SPMV_AUTO)
    ref = orc.spmv(orc.CSR(n, A.ai, A.aj, A.av), x.cpu().numpy())
    rel = float(np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref))
    if rel > 1e-12:
        raise SystemExit("SpMV parity failure: rel err %.3e" % rel)

    durs, wall_ms = time_spmv(la, torch, mats, x, y, args.steps, args.warmup, la.SPMV_AUTO,
                              la.F64, la.F64, stream)
    B = spmv_bytes(n, nnz)
    avg_ms = float(np.mean(durs))
    gbs_kernel = B / (avg_ms * 1e-3) / 1e9
    gbs_wall = B * args.steps / (wall_ms * 1e-3) / 1e9
    extra = {}
    # the other SpMV forms on the same matrix, for the record
    for name, alg, val, acc in [("albus_f64", la.SPMV_ALBUS, la.F64, la.F64),
                                ("row_f64", la.SPMV_ROW, la.F64, la.F64)]:
        d2, _ = time_spmv(la, torch, mats, x, y, 20, 3, alg, val, acc, stream)
        extra[name] = round(B / (np.mean(d2) * 1e-3) / 1e9, 1)
    x32 = x.float()
    y32 = torch.zeros(n, dtype=torch.float32, device=dev)
    for dt, sv, name in [(la.F32, 4, "f32"), (la.BF16, 2, "bf16")]:
        for m in mats:
            m.add_dtype(dt)
        d2, _ = time_spmv(la, torch, mats, x32, y32, 20, 3, la.SPMV_AUTO, dt, la.F32, stream)
        Bl = spmv_bytes(n, nnz, s_v=sv, s_x=4, s_y=4)
        extra["%s_%s_vals_f32_vecs" % (ALG_NAMES[chosen], name)] = {
            "GBs": round(Bl / (np.mean(d2) * 1e-3) / 1e9, 1),
            "us": round(float(np.mean(d2)) * 1e3, 2)}

    # the bytes the timed kernel's own layout streams (for sliced ELL: every
    # slot's value and column, padding included, the slice table, x once
    # and y), next to the CSR figure `value` is quoted in
    layout = None
    if chosen == 4 and sinfo["built"]:
        nsl = (n + 63) // 64
        Bl = sinfo["slots"] * (8 + (2 if sinfo["col16"] else 4)) + nsl * (8 + 4) + n * 8 + n * 8
        layout = {"name": "sliced ELL, %s columns" % ("16-bit offset" if sinfo["col16"]
                                                      else "int32"),
                  "bytes_per_spmv": int(Bl), "GBs": round(Bl / (avg_ms * 1e-3) / 1e9, 1),
                  "frac": round(Bl / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    copy = copy_ceiling(torch, dev)
    if layout:
        layout["frac_of_copy"] = round(layout["GBs"] / copy["GBs"], 4)
        layout["frac_of_read"] = round(layout["GBs"] / copy["read"]["GBs"], 4)
    traffic = None
    tpath = os.path.join(ROOT, "profiles",
                         "traffic_spmv_%s_f64_laplace3d_128.json" % ALG_NAMES[chosen])
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    pcg = None
    if not args.no_pcg:
        del mats[1:]
        mode = la.TRSV_LEVEL if args.trsv == "level" else la.TRSV_SYNCFREE
        variants = [("fp64", dict(precision=la.PREC_FP64, stop_rule=la.STOP_TRUE_RESIDUAL)),
                    ("fp64_recurrence", dict(precision=la.PREC_FP64,
                                             stop_rule=la.STOP_RECURRENCE)),
                    ("mixed_f32", dict(precision=la.PREC_MIXED_F32)),
                    ("mixed_bf16", dict(precision=la.PREC_MIXED_BF16))]
        for k in (3, 1):  # Jacobi sweeps per triangle
            J = dict(trsv_mode=la.TRSV_JACOBI, sweeps=k)
            variants += [("fp64_jacobi%d" % k, dict(precision=la.PREC_FP64,
                                                    stop_rule=la.STOP_TRUE_RESIDUAL, **J)),
                         ("fp64_jacobi%d_recurrence" % k, dict(precision=la.PREC_FP64,
                                                               stop_rule=la.STOP_RECURRENCE,
                                                               **J)),
                         ("mixed_f32_jacobi%d" % k, dict(precision=la.PREC_MIXED_F32, **J)),
                         ("mixed_bf16_jacobi%d" % k, dict(precision=la.PREC_MIXED_BF16, **J))]
        pcg = run_pcg(la, torch, A, dev, stream, mode, variants, cpu=not args.no_cpu)
        r = pcg["runs"]
        sec = {k: v["seconds"] for k, v in r.items()}
        # the headline: mixed precision against FP64 with the same
        # preconditioner (same triangular solves), FP64 at its faster stop
        # rule; the exact IC(0) apply is the reference's preconditioner
        groups = {}
        for tag, label in [("", "exact IC(0) (%s)" % args.trsv), ("_jacobi3", "3 Jacobi sweeps"),
                           ("_jacobi1", "1 Jacobi sweep")]:
            f = "fp64" + tag
            fr = "fp64%s_recurrence" % tag
            ref = f if sec[f] <= sec[fr] else fr
            groups[label] = (ref, ["mixed_f32" + tag, "mixed_bf16" + tag])
        pcg["mixed_vs_fp64_same_preconditioner"] = mixed_summary(sec, groups)
        # for the record only: the Jacobi variants change the preconditioner
        # (and the iteration count), so these are not precision speedups
        pcg["ratio_to_plain_fp64_exact_ic0"] = {k: round(sec["fp64"] / sec[k], 3)
                                                for k in sec if k.startswith("mixed")}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(A, args.cpu_seconds)

    # the other configurations' records, each on its own (c2's matrices and
    # plans freed first)
    sub = {}
    if not args.no_extra:
        del mats
        import gc
        gc.collect()
        torch.cuda.empty_cache()
        for name, fn in [("c4", bench_c4), ("c3", bench_c3), ("c5", bench_c5)]:
            t0 = time.time()
            sub[name] = fn(la, torch, dev, stream, args)
            sub[name]["record_wall_s"] = round(time.time() - t0, 1)
            gc.collect()
            torch.cuda.empty_cache()
            log("record", name, "done in %.1f s" % sub[name]["record_wall_s"])

    line = {
        "metric": METRIC,
        "value": round(gbs_kernel, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3-D 7-point Laplacian 128^3, FP64 CSR int32, RCM-ordered; "
                               "y = A x (%s)" % ALG_DESC[chosen],
                   "n": n, "nnz": nnz, "bytes_per_spmv": B,
                   "sell": {"col16": sinfo["col16"],
                            "fill": round(sinfo["slots"] / nnz, 4) if sinfo["built"] else None},
                   "pct_hbm_peak": round(100 * gbs_kernel / HBM_PEAK_GBS, 1),
                   "wall_GBs": round(gbs_wall, 1), "host_prep": prep,
                   "other_spmv": extra},
        "roofline": {"bound": "hbm", "achieved": round(gbs_kernel, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs_kernel / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": (ALG_KERNEL[chosen] % ("short" if sinfo["col16"] else "int")
                                if chosen == 4 else ALG_KERNEL[chosen]),
                     "avg_launch_us": round(avg_ms * 1e3, 3),
                     "achieved_basis": "CSR algorithmic bytes (nnz*(8+4) + (n+1)*4 + n*8 + "
                                       "n*8) over the measured launch duration",
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this "
                                        "kernel on this matrix, scaled by the factors that "
                                        "make a 1 GiB device copy read and write 1 GiB "
                                        "(committed counter summary %s, not measured by this "
                                        "run)" % os.path.relpath(tpath, ROOT)) if traffic else None,
                     "traffic_GBs": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "traffic_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                     if traffic else None,
                     "layout": layout,
                     "copy_ceiling": copy},
        "cpu_baseline": cpu,
        "pcg": pcg,
    }
    line.update(sub)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

"""Benchmark: cell-updates/s and HBM roofline fraction of the OPS hot path on one MI355X.

Default workload: CloverLeaf 2D, deck clover_bm16_short.in (3840 x 3840 cells, 87 hydro
steps, FP64), through libops_hip.so and the translated application. A step is one hydro
time step of the whole mesh; a cell-update is one cell advanced by one step.

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-laplace] [--no-tealeaf]
                    [--tea-cells N] [--tea-steps S] [--tea-solvers 0,1,3,4] [--no-learned]
                    [--unet-pairs P] [--unet-epochs E] [--no-adi] [--adi-steps S]
                    [--no-clover3d]

Prints one JSON line. `value` is measured with the inputs resident in HBM: the deck is
generated on the device before the timed region; the K timed steps are bracketed by
device synchronisation. Per-kernel durations come from HIP events recorded on the
runtime's stream around every launch inside the timed region; `roofline` is computed for
the kernel with the largest total time from its ops_compute_transfer bytes per launch.
`cpu_baseline` runs the same application on the oracle's host backend (OpenMP, one thread
per physical core of the process's CPU share, bound close, lscpu model recorded) for a
bounded sample: one warm-up, then the median of three runs; each other workload carries its
own `cpu_baseline` the same way. `laplace` is the 2D Laplace Jacobi workload (1024^2, 1000
iterations). `tealeaf` reports the second half of the metric:
TeaLeaf 4096^2 time-to-solution of plain FP64 CG against the multigrid-preconditioned
solver (FP64 outer CG, FP32 V-cycle), with residuals and field differences;
`tealeaf_learned` the learned warm start (data generation, U-Net training throughput
against the BF16 matrix-core peak, iterations and time with and without it); `adi` the
batched tridiagonal application.
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cell-updates/s + % HBM roofline; TeaLeaf time-to-tol learned+mixed vs FP64 CG"
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
BF16_MFMA_PEAK_TFLOPS = 2500.0  # MI355X published dense BF16 matrix throughput
GOLD_KE_BM16 = 0.307475452287895  # apps/c/CloverLeaf/field_summary.cpp:94

BM16_DECK = """*clover
 state 1 density=0.2 energy 1.0
 state 2 density=1.0 energy=2.5 geometry=rectangle xmin=0.0 xmax=5.0 ymin=0.0 ymax=2.0
 x_cells={n}
 y_cells={n}
 xmin=0.0
 ymin=0.0
 xmax=10.0
 ymax=10.0
 initial_timestep=0.04
 timestep_rise=1.5
 max_timestep=0.04
 end_time=0.55
 end_step=87
 test_problem 4
*endclover
"""


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profile(kernel, algorithmic=None):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    committed under profiles/ (null when not measured for this kernel). The counters are
    the median over a profiled run's dispatches, so a kernel launched on meshes of several
    sizes reports its most frequent size: when that value is not within 0.5-4x of this
    launch's algorithmic bytes it was measured at another size, and null is returned."""
    p = ROOT / "profiles" / "pmc_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
        e = d.get("kernels", {}).get(kernel)
        if e is None:
            return None
        t = float(e["bytes_per_launch"])
        if algorithmic and not 0.5 <= t / algorithmic <= 4.0:
            return None
        return t
    except Exception:
        return None


def kernel_short_name(name):
    """A demangled kernel name without its return type, namespaces, ops_kern_ prefix and
    parameter list; template arguments kept (trid_part<double, false, true, false>)."""
    name = name.replace("(anonymous namespace)::", "")
    if name.startswith("void "):
        name = name[5:]
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):  # the parameter list: the first '(' outside <...>
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    short = name[:cut]
    head, _, tail = short.partition("<")
    short = head.split("::")[-1] + (("<" + tail) if tail else "")
    if short.startswith("ops_kern_"):
        short = short[len("ops_kern_"):]
    return short


def ingest_pmc(d, match=None):
    """rocprofv3 --pmc CSVs (one run with FETCH_SIZE, one with WRITE_SIZE; values in KiB per
    dispatch) -> profiles/pmc_traffic.json. Each counter is scaled by what it reads for the
    runtime's copy probe kernel, whose true traffic is known (n bytes read, n written).
    Kernels measured here replace their entries; the file's other kernels are kept. With
    `match` (a regular expression), only kernels whose full name it finds are taken (the copy
    probe always is)."""
    import csv
    import glob
    import re
    rows = []
    for f in glob.glob(str(d / "**" / "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            rows += list(csv.DictReader(fh))
    if match:
        pat = re.compile(match)
        rows = [r for r in rows if "copy_probe_kernel" in r["Kernel_Name"] or pat.search(r["Kernel_Name"])]
    per = {}
    for r in rows:
        short = kernel_short_name(r["Kernel_Name"])
        per.setdefault((short, r["Counter_Name"]), []).append(float(r["Counter_Value"]) * 1024.0)
    probe_bytes = float(1 << 30)
    out = {"source": str(d), "unit": "bytes per dispatch", "kernels": {}}
    cal = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        v = per.get(("copy_probe_kernel", c))
        if v:
            v = sorted(v)[len(v) // 2]
            cal[c] = v / probe_bytes
    out["calibration"] = {"kernel": "copy_probe_kernel", "bytes_each_way": probe_bytes,
                          "counter_over_true": cal}
    names = {k for k, _ in per}
    for k in sorted(names):
        f, w = per.get((k, "FETCH_SIZE")), per.get((k, "WRITE_SIZE"))
        if not f or not w or k == "copy_probe_kernel":
            continue
        fm, wm = sorted(f)[len(f) // 2], sorted(w)[len(w) // 2]
        fs, ws = cal.get("FETCH_SIZE", 1.0), cal.get("WRITE_SIZE", 1.0)
        out["kernels"][k] = {"fetch_bytes": fm / fs, "write_bytes": wm / ws,
                             "bytes_per_launch": fm / fs + wm / ws, "dispatches": len(f)}
    p = ROOT / "profiles" / "pmc_traffic.json"
    if p.exists():
        old = json.loads(p.read_text())
        merged = dict(old.get("kernels", {}))
        merged.update(out["kernels"])
        srcs = [x for x in old.get("sources", [old["source"]] if "source" in old else []) if x != str(d)]
        out = {"sources": srcs + [str(d)], "unit": out["unit"], "calibration": out["calibration"],
               "kernels": dict(sorted(merged.items()))}
    p.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {p} ({len(out['kernels'])} kernels, calibration {cal})")


def roofline(stats):
    k = max(stats, key=lambda n: stats[n]["ms"])
    e = stats[k]
    per_launch_bytes = e["bytes"] / e["calls"]
    avg_ms = e["ms"] / e["calls"]
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    return k, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
               "traffic": traffic_from_profile(k, per_launch_bytes), "kernel": k,
               "bytes_per_launch": per_launch_bytes, "avg_launch_us": round(avg_ms * 1e3, 2)}


def top_kernels(stats, k):
    """The k kernels with the most device time: share of time, calls, GB/s (algorithmic)."""
    tot = sum(e["ms"] for e in stats.values()) or 1.0
    rows = sorted(stats.items(), key=lambda kv: -kv[1]["ms"])[:k]
    return [{"kernel": n, "time_frac": round(e["ms"] / tot, 3), "calls": e["calls"],
             "avg_us": round(e["ms"] * 1e3 / e["calls"], 2),
             "GBps": round(e["bytes"] / (e["ms"] * 1e-3) / 1e9, 1) if e["ms"] else None}
            for n, e in rows]


def whole_step_bw(stats, seconds):
    tot_bytes = sum(e["bytes"] for e in stats.values())
    tot_ms = sum(e["ms"] for e in stats.values())
    return {"algorithmic_GB": round(tot_bytes / 1e9, 3),
            "kernel_busy_s": round(tot_ms * 1e-3, 4),
            "GBps_over_kernel_time": round(tot_bytes / (tot_ms * 1e-3) / 1e9, 1) if tot_ms else None,
            "GBps_over_wall": round(tot_bytes / seconds / 1e9, 1)}


_HOST_CPU = None


def host_cpu():
    """The host CPU the baselines run on: model, sockets and physical cores from lscpu; the
    OpenMP threads used are one per physical core of this process's CPU share (at most the
    OMP_NUM_THREADS the environment sets), bound close to one core each (OMP_PROC_BIND=close,
    OMP_PLACES=cores, set before the OpenMP runtime of the oracle libraries loads). Computed
    once: after the first OpenMP region the binding pins the main thread to one core, so the
    process's affinity no longer shows its CPU share."""
    global _HOST_CPU
    if _HOST_CPU is not None:
        return _HOST_CPU
    import subprocess
    info = {"model": None, "sockets": None, "physical_cores": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v.strip())
        cores = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                               timeout=10).stdout
        info["physical_cores"] = len({l for l in cores.splitlines() if l and not l.startswith("#")})
    except Exception:
        pass
    share = len(os.sched_getaffinity(0))
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    phys = info["physical_cores"] or share
    threads = max(1, min(phys, share, cap))
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    os.environ["OMP_NUM_THREADS"] = str(threads)
    info.update({"threads": threads, "cpu_share": share, "OMP_PROC_BIND": "close",
                 "OMP_PLACES": "cores"})
    _HOST_CPU = info
    return info


def median_time(fn, runs=3):
    """One warm-up call, then the median wall time of `runs` calls."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_lib(app, threads):
    """The oracle's host build of an application (the CPU baseline: a port, OpenMP)."""
    from ops_amd.apps import SIGNATURES
    lib = ctypes.CDLL(str(ROOT / "oracle" / "_build" / f"lib{app}_cpu.so"), mode=ctypes.RTLD_LOCAL)
    for fn, (res, argt) in SIGNATURES[app].items():
        if hasattr(lib, fn):
            getattr(lib, fn).restype = res
            getattr(lib, fn).argtypes = argt
    lib.ops_cpu_set_threads.argtypes = [ctypes.c_int]
    lib.ops_cpu_set_threads(threads)
    return lib


def laplace_gpu(args):
    """Laplace 2D Jacobi, 1024^2 interior, 1000 iterations of the reference's loop: the
    5-point stencil with its MAX reduction, ops_reduction_result (read on the host every
    iteration, as the reference's stop test does), the copy back. 32 algorithmic bytes per
    cell-update (16 stencil + 16 copy)."""
    import ops_amd
    L = ops_amd.app("laplace2d")
    R = ops_amd.runtime()
    n, iters = args.lap_cells, args.lap_iters
    argc, argv = ops_amd.argv()
    err = (ctypes.c_double * 1)()
    assert L.laplace2d_init(argc, argv, n, n, 0) == 0
    assert L.laplace2d_iterate(20, err) == 0  # warm-up (autotuning candidates)
    assert L.laplace2d_finalize() == 0
    assert L.laplace2d_init(argc, argv, n, n, 0) == 0
    R.ops_hip_device_sync()
    t0 = time.perf_counter()
    assert L.laplace2d_iterate(iters, err) == 0
    R.ops_hip_device_sync()
    t1 = time.perf_counter()
    final_err = err[0]
    assert L.laplace2d_finalize() == 0
    # per-kernel durations from an eager pass of the same iterations with HIP events
    assert L.laplace2d_init(argc, argv, n, n, 0) == 0
    R.ops_hip_device_sync()
    R.ops_hip_reset_stats()
    R.ops_hip_set_timing(1)
    assert L.laplace2d_iterate(iters, err) == 0
    R.ops_hip_device_sync()
    R.ops_hip_set_timing(0)
    assert L.laplace2d_finalize() == 0
    stats = {k: v for k, v in ops_amd.kernel_stats().items() if k in ("apply_stencil", "copy")}
    secs = t1 - t0
    _, roof = roofline({"apply_stencil": stats["apply_stencil"]})
    return {"workload": f"laplace2d {n}x{n}, {iters} Jacobi iterations (stencil + MAX reduction + copy)",
            "value": n * n * iters / secs, "unit": "cell-updates/s", "s": round(secs, 4),
            "us_per_iteration": round(secs * 1e6 / iters, 2),
            "algorithmic_GBps_over_wall": round(32.0 * n * n * iters / secs / 1e9, 1),
            "roofline": roof, "top_kernels": top_kernels(stats, 2), "final_error": final_err}


def laplace_cpu(args, cpu):
    import numpy as np
    import oracle
    n, iters = args.lap_cells, args.lap_cpu_iters
    A = np.zeros((n + 2, n + 2))
    An = np.zeros_like(A)
    pa, pn = A.ctypes.data_as(ctypes.c_void_p), An.ctypes.data_as(ctypes.c_void_p)
    t = median_time(lambda: oracle.lib().oracle_laplace(n, n, iters, 0, pa, pn, cpu["threads"]))
    return {"value": n * n * iters / t, "unit": "cell-updates/s", "cores": cpu["threads"],
            "kind": "port", "cpu": cpu,
            "sample": f"Laplace {n}x{n}, {iters} iterations of the oracle's C restatement "
                      f"(OpenMP), median of 3 runs after a warm-up: {t:.3f} s"}


def clover_gpu(args, R):
    import ops_amd
    L = ops_amd.app("cloverleaf")
    n = args.cells
    with tempfile.TemporaryDirectory() as t:
        deck = Path(t) / "clover.in"
        deck.write_text(BM16_DECK.format(n=n))
        argc, argv = ops_amd.argv()
        if args.warmup > 0:
            log(f"[bench] warmup: {args.warmup} steps")
            assert L.cloverleaf_init(argc, argv, str(deck).encode()) == 0
            L.cloverleaf_set_chain(args.chain)
            assert L.cloverleaf_step(args.warmup) == 0
            assert L.cloverleaf_finalize() == 0
        log("[bench] generating the deck on the device")
        assert L.cloverleaf_init(argc, argv, str(deck).encode()) == 0
        L.cloverleaf_set_chain(args.chain)
        # measured HBM bandwidth: a plain 1 GiB device copy (also the counter calibration)
        checks_copy = R.ops_hip_copy_probe(1 << 30, 10)
        graph = not args.no_graph
        if graph:
            # capture the two step graphs before the timed region (dt stays on the GPU)
            L.cloverleaf_set_graph(1)
            assert L.cloverleaf_capture() == 0
        R.ops_hip_device_sync()
        log(f"[bench] timed: {args.steps} steps ({'captured' if graph else 'eager'})")
        if not graph:
            R.ops_hip_reset_stats()
            R.ops_hip_set_timing(1)
        t0 = time.perf_counter()
        rc = L.cloverleaf_step(args.steps)
        R.ops_hip_device_sync()
        t1 = time.perf_counter()
        R.ops_hip_set_timing(0)
        L.cloverleaf_set_graph(0)
        assert rc == 0
        s = (ctypes.c_double * 5)()
        assert L.cloverleaf_summary(s) == 0
        assert L.cloverleaf_finalize() == 0
        eager_s = None
        if graph:
            # per-kernel durations (HIP events around every launch) from an eager pass of
            # the same steps: the replayed graphs run the same kernels
            log(f"[bench] eager pass with per-kernel timing: {args.steps} steps")
            assert L.cloverleaf_init(argc, argv, str(deck).encode()) == 0
            L.cloverleaf_set_chain(args.chain)
            R.ops_hip_device_sync()
            R.ops_hip_reset_stats()
            R.ops_hip_set_timing(1)
            e0 = time.perf_counter()
            assert L.cloverleaf_step(args.steps) == 0
            R.ops_hip_device_sync()
            eager_s = time.perf_counter() - e0
            R.ops_hip_set_timing(0)
            assert L.cloverleaf_finalize() == 0
        stats = ops_amd.kernel_stats()
    secs = t1 - t0
    checks = {"ke": s[3], "hbm_copy_GBps": round(checks_copy, 1), "captured_steps": graph,
              "loop_chain_bands": args.chain}
    if eager_s is not None:
        checks["eager_ms_per_step_with_kernel_timing"] = round(eager_s * 1e3 / args.steps, 3)
    if args.steps == 87 and n == 3840:
        checks["ke_golden"] = GOLD_KE_BM16
        checks["ke_pct_diff"] = abs(100.0 * (s[3] / GOLD_KE_BM16) - 100.0)
        checks["ke_pass_0.001pct"] = checks["ke_pct_diff"] < 0.001
    return secs, stats, checks


def tealeaf_gpu(args):
    """TeaLeaf 4096^2, 10 implicit steps to eps = 1e-10 (the reference's stop rule on the
    CG residual): plain FP64 CG through the generated loops (the reference algorithm on
    this backend) against the multigrid-preconditioned solver. A third run to eps = 1e-12
    gives a tighter solution to measure both fields' errors against."""
    import numpy as np
    import ops_amd
    from ops_amd.apps.tealeaf.decks import tea_deck
    L = ops_amd.app("tealeaf")
    R = ops_amd.runtime()
    n, steps = args.tea_cells, args.tea_steps

    def run(solver, eps):
        with tempfile.TemporaryDirectory() as t:
            deck = Path(t) / "tea.in"
            deck.write_text(tea_deck(n, steps, eps=eps, max_iters=20000, extra=("tl_check_result",)))
            argc, argv = ops_amd.argv()
            assert L.tealeaf_init(argc, argv, str(deck).encode()) == 0
            assert L.tealeaf_set_solver(solver) == 0
            R.ops_hip_device_sync()
            st = np.zeros(7)
            iters, exact, rel = [], [], []
            t0 = time.perf_counter()
            for _ in range(steps):
                assert L.tealeaf_step(1) == 0
                L.tealeaf_state(st.ctypes.data_as(ctypes.c_void_p))
                iters.append(int(st[2]))
                exact.append(float(st[6]))
                rel.append(float(st[3] / st[4]))
            R.ops_hip_device_sync()
            t1 = time.perf_counter()
            cnt = L.tealeaf_fetch(b"u", None)
            u = np.zeros(cnt)
            L.tealeaf_fetch(b"u", u.ctypes.data_as(ctypes.c_void_p))
            s4 = np.zeros(4)
            assert L.tealeaf_summary(s4.ctypes.data_as(ctypes.c_void_p)) == 0
            assert L.tealeaf_finalize() == 0
        log(f"[bench] tealeaf solver {solver} eps {eps}: {t1 - t0:.3f} s, {sum(iters)} iterations")
        return {"s": t1 - t0, "iters": iters, "exact": exact, "rel": rel, "temp": s4[3], "u": u}

    names = {0: "fp64_cg_plain", 1: "fp64_cg_device", 2: "fp32_cg_fp64_refinement",
             3: "fp64_pcg_fp32_multigrid", 4: "fp64_pcg_fp64_multigrid"}
    runs = {}
    for solver in args.tea_solvers:
        runs[solver] = run(solver, 1e-10)
    tight = run(3, 1e-12)

    def maxrel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    out = {
        "workload": f"tealeaf_2d {n}x{n}, {steps} steps, 8 seeded rectangles, tl_eps 1e-10",
        "time_to_solution_s": {names[k]: round(v["s"], 4) for k, v in runs.items()},
        "iterations": {names[k]: sum(v["iters"]) for k, v in runs.items()},
        "max_relative_cg_residual": {names[k]: max(v["rel"]) for k, v in runs.items()},
        "final_exact_residual": {names[k]: v["exact"][-1] for k, v in runs.items()},
        "field_max_rel_err_vs_eps1e-12_solution": {names[k]: maxrel(v["u"], tight["u"]) for k, v in runs.items()},
        "field_target": "1e-9 of plain FP64 CG; plain FP64 CG itself stops at the error above from "
                        "the converged (eps 1e-12) solution",
        "temperature": {names[k]: v["temp"] for k, v in runs.items()},
    }
    if 0 in runs:
        out["field_max_rel_diff_vs_fp64_cg"] = {names[k]: maxrel(v["u"], runs[0]["u"])
                                               for k, v in runs.items() if k != 0}
        out["speedup_vs_fp64_cg_plain"] = {names[k]: round(runs[0]["s"] / v["s"], 2)
                                           for k, v in runs.items() if k != 0}
    if 1 in runs:
        out["speedup_vs_fp64_cg_device"] = {names[k]: round(runs[1]["s"] / v["s"], 2)
                                            for k, v in runs.items() if k not in (0, 1)}
    if 3 in runs and 4 in runs:
        # what FP32 storage and arithmetic in the V-cycle buy: the same preconditioned solve
        # with the V-cycle in FP64
        out["fp32_vcycle_gain"] = {"time_ratio_fp64_over_fp32": round(runs[4]["s"] / runs[3]["s"], 3),
                                   "iterations_fp32": sum(runs[3]["iters"]),
                                   "iterations_fp64": sum(runs[4]["iters"])}
    return out


def tealeaf_cpu(args, cpu):
    """Plain FP64 CG of the oracle's host build on the 4096^2 deck, time per CG iteration:
    the first implicit step solved with the iteration cap at 1 and at 1 + --tea-cpu-iters,
    the difference of the two step times over the extra iterations (the step's set-up
    cancels); one warm-up pair, then the median of three. The time-to-solution beside it is
    that time the GPU run's plain-CG iteration count (an extrapolation)."""
    from ops_amd.apps.tealeaf.decks import tea_deck
    import ops_amd
    L = cpu_lib("tealeaf", cpu["threads"])
    n, its = args.tea_cells, args.tea_cpu_iters
    argc, argv = ops_amd.argv()

    def step_time(cap):
        with tempfile.TemporaryDirectory() as t:
            deck = Path(t) / "tea.in"
            deck.write_text(tea_deck(n, 1, eps=1e-10, max_iters=cap))
            assert L.tealeaf_init(argc, argv, str(deck).encode()) == 0
            assert L.tealeaf_set_solver(0) == 0
            t0 = time.perf_counter()
            assert L.tealeaf_step(1) == 0
            dt = time.perf_counter() - t0
            assert L.tealeaf_finalize() == 0
        return dt

    diffs = []
    for k in range(4):
        d = (step_time(1 + its) - step_time(1)) / its
        if k:
            diffs.append(d)
    per = sorted(diffs)[1]
    return {"value": n * n / per, "unit": "cell-iterations/s", "cores": cpu["threads"],
            "kind": "port", "cpu": cpu, "s_per_cg_iteration": per,
            "sample": f"TeaLeaf {n}x{n}, plain FP64 CG on the oracle host build (OpenMP): the "
                      f"first implicit step capped at 1 and at {1 + its} iterations, difference "
                      f"per iteration, median of 3 after a warm-up: {per * 1e3:.2f} ms per "
                      f"iteration"}


def unet_training_gpu(args):
    """U-Net training throughput alone (no TeaLeaf solves): --unet-synthetic random samples
    at 1024^2 (4 features uniform on [-1, 1], targets normal), Adam with batch 4, one
    warm-up epoch, then two timed epochs; TFLOP/s by the network's training flop count.
    For profiling the training kernels in a short run."""
    import ctypes
    import numpy as np
    import ops_amd
    from ops_amd.apps.tealeaf import learned
    L = ops_amd.app("tealeaf")
    net, cnt = 1024, args.unet_synthetic
    assert L.tea_unet_init(7) == 0
    assert L.tea_unet_dataset_reserve(cnt, net, net) == 0
    rng = np.random.default_rng(1)
    for i in range(cnt):
        x = rng.uniform(-1, 1, (net, net, 4)).astype(np.float32)
        t = rng.normal(0, 1, (net, net)).astype(np.float32)
        assert L.tea_unet_dataset_set(i, x.ctypes.data_as(ctypes.c_void_p),
                                      t.ctypes.data_as(ctypes.c_void_p)) == 0
    learned.train(L, 1, batch=4)
    loss, s = learned.train(L, 2, batch=4)
    assert L.tea_unet_free() == 0
    flops = learned.training_flops_per_pixel() * net * net * cnt * 2
    return {"workload": f"U-Net training, {cnt} synthetic samples at {net}^2, batch 4, 2 epochs",
            "s": round(s, 4), "ms_per_sample": round(s * 1e3 / (2 * cnt), 3),
            "TFLOPs": round(flops / s / 1e12, 2),
            "frac_of_bf16_mfma_peak": round(flops / s / 1e12 / BF16_MFMA_PEAK_TFLOPS, 5)}


def tealeaf_learned_gpu(args):
    """Learned warm start and preconditioner correction: training pairs at 1024^2 from
    seeded rectangle decks (seeds 20261203 + k for k >= 1, so the benchmark's own layout,
    seed 20261203, is not among them), each deck run for --unet-steps-per-deck implicit
    steps with the FP64 multigrid PCG to 1e-10; Adam training of the U-Net on the GPU; then
    the 4096^2 benchmark solve (multigrid PCG, --tea-steps steps, tl_eps 1e-10) without the
    network and with it in four modes, the network evaluated on 4 x 4 averages:
      1  warm start kept only when its residual is below the plain start's,
      2  warm start always kept (its energy-norm error never exceeds the plain start's),
      3  warm start always kept and the network's change added to the multigrid
         preconditioner as a rank-one A-orthogonal correction,
      4  that correction alone, from the plain start.
    The stop threshold is the plain start's in every mode. `best_case` bounds what any
    warm start can gain: each step started from its converged solution with a smooth
    relative error e."""
    import numpy as np
    import ops_amd
    from ops_amd.apps.tealeaf import learned
    from ops_amd.apps.tealeaf.decks import tea_deck
    L = ops_amd.app("tealeaf")
    net = 1024
    assert L.tea_unet_init(7) == 0
    pairs, epochs = args.unet_pairs, args.unet_epochs
    gen_s = learned.generate(L, pairs, cells=net, steps_per_deck=args.unet_steps_per_deck, seed0=20261204)
    log(f"[bench] learned: {pairs} pairs at {net}^2 in {gen_s:.1f} s")
    loss, train_s = learned.train(L, epochs, batch=4, lr=1e-3, seed=1)
    flops = learned.training_flops_per_pixel() * net * net * pairs * epochs
    log(f"[bench] learned: {epochs} epochs in {train_s:.1f} s, loss {loss}")
    n, steps = args.tea_cells, args.tea_steps
    deck = tea_deck(n, steps, eps=1e-10, max_iters=20000)
    plain = learned.solve_run(L, deck, 3, steps, fetch=("u",))
    runs = {m: learned.solve_run(L, deck, 3, steps, net_cells=net, fetch=("u",), mode=m)
            for m in (1, 2, 3, 4)}
    assert L.tea_unet_free() == 0
    names = {1: "warm_start_if_better", 2: "warm_start", 3: "warm_start_and_correction",
             4: "correction"}
    diff = {names[m]: float(np.max(np.abs(o["u"] - plain["u"])) / np.max(np.abs(plain["u"])))
            for m, o in runs.items()}
    saved = {names[m]: sum(plain["iters"]) - sum(o["iters"]) for m, o in runs.items()}
    for m, o in runs.items():
        log(f"[bench] learned: mode {m} iterations {sum(o['iters'])} (plain {sum(plain['iters'])}), "
            f"{o['s']:.3f} s (plain {plain['s']:.3f} s)")
    best = learned.best_case(L, deck, 3, steps, errors=(0.1, 0.01))
    log(f"[bench] learned: best case {dict((str(k), sum(v)) for k, v in best.items())}")
    return {
        "workload": f"tealeaf_2d {n}x{n}, {steps} steps, tl_eps 1e-10, FP64 multigrid PCG; "
                    f"U-Net trained on {pairs} pairs at {net}^2, {epochs} epochs",
        "data_generation_s": round(gen_s, 2),
        "training": {"s": round(train_s, 2), "loss_per_epoch": [float(v) for v in loss],
                     "TFLOPs": round(flops / train_s / 1e12, 2),
                     "frac_of_bf16_mfma_peak": round(flops / train_s / 1e12 / BF16_MFMA_PEAK_TFLOPS, 5)},
        "iterations": {"plain": plain["iters"],
                       **{names[m]: o["iters"] for m, o in runs.items()},
                       "saved": saved[names[3]], "saved_by_mode": saved},
        "warm_start_kept": runs[1]["kept"],
        "initial_residual": {"plain": plain["initial"], "warm_start": runs[2]["warm"]},
        "time_to_solution_s": {"plain": round(plain["s"], 4),
                               **{names[m]: round(o["s"], 4) for m, o in runs.items()}},
        "field_max_rel_diff": diff,
        "best_case": {"what": "each step started from its converged solution plus a smooth "
                              "error of relative size e (2-norm of the step's change)",
                      "iterations": {("plain" if k == "plain" else f"e={k:g}"): sum(v)
                                     for k, v in best.items()},
                      "per_step": {("plain" if k == "plain" else f"e={k:g}"): v
                                   for k, v in best.items()}},
    }


def adi_gpu(args):
    """Batched ADI: 48 blocks with extents in {64..256} (splitmix64 seed 5), u ~ U[0,1),
    lambda 1, 100 steps of preproc + x, y, z tridiagonal solves, one step replayed as a
    graph. Algorithmic bytes per cell-step: preproc 8 + 80, x and y solves 40 each, z solve
    with increment 48 (216 B)."""
    import numpy as np
    import ops_amd
    from ops_amd.apps.adi.config import adi_sizes
    L = ops_amd.app("adi")
    R = ops_amd.runtime()
    sizes = adi_sizes()
    cells = sum(a * b * c for a, b, c in sizes)
    sz = np.array([v for s in sizes for v in s], dtype=np.int32)
    argc, argv = ops_amd.argv()
    # warm-up session with eager launches (autotuning happens here), then the timed one
    os.environ["OPS_ADI_GRAPH"] = "0"
    assert L.adi_init(argc, argv, len(sizes), sz.ctypes.data_as(ctypes.c_void_p), 1, 5, 1.0) == 0
    assert L.adi_step(3) == 0
    assert L.adi_finalize() == 0
    os.environ["OPS_ADI_GRAPH"] = "0" if args.adi_eager else "1"
    assert L.adi_init(argc, argv, len(sizes), sz.ctypes.data_as(ctypes.c_void_p), 1, 5, 1.0) == 0
    assert L.adi_step(2) == 0  # eager step + graph capture + one replay
    R.ops_hip_device_sync()
    t0 = time.perf_counter()
    assert L.adi_step(args.adi_steps) == 0
    R.ops_hip_device_sync()
    t1 = time.perf_counter()
    assert L.adi_finalize() == 0
    secs = t1 - t0
    gbs = cells * 216.0 * args.adi_steps / secs / 1e9
    return {"workload": f"adi_3d 48 blocks, {cells} cells, {args.adi_steps} steps",
            "s": round(secs, 4), "ms_per_step": round(secs * 1e3 / args.adi_steps, 3),
            "cell_steps_per_s": cells * args.adi_steps / secs,
            "algorithmic_GBps": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}


def adi_cpu(args, cpu):
    """The oracle host build of the ADI application on the first --adi-cpu-blocks blocks of
    the 48-block configuration (the same generator and extents), one step timed."""
    import numpy as np
    from ops_amd.apps.adi.config import adi_sizes
    import ops_amd
    L = cpu_lib("adi", cpu["threads"])
    sizes = adi_sizes()[:args.adi_cpu_blocks]
    cells = sum(a * b * c for a, b, c in sizes)
    sz = np.array([v for s_ in sizes for v in s_], dtype=np.int32)
    argc, argv = ops_amd.argv()
    assert L.adi_init(argc, argv, len(sizes), sz.ctypes.data_as(ctypes.c_void_p), 1, 5, 1.0) == 0
    t = median_time(lambda: L.adi_step(1))
    assert L.adi_finalize() == 0
    return {"value": cells / t, "unit": "cell-steps/s", "cores": cpu["threads"], "kind": "port",
            "cpu": cpu, "sample": f"ADI, blocks 0..{len(sizes) - 1} of the 48 ({cells} cells), one "
                                  f"step on the oracle host build (OpenMP), median of 3 after a "
                                  f"warm-up: {t:.3f} s"}


def clover3d_gpu(args, R):
    """CloverLeaf 3D (apps/c/CloverLeaf_3D, an API v1 application) on the reference's deck
    clover.in (96^3, 87 steps, known answer ke 20.0546870878964) and on a larger mesh of
    the same problem (--c3-cells^3, --c3-steps steps) that fills the GPU. A warm-up session
    runs first (the autotuner picks workgroup shapes there); the timed session starts from
    the generated state with every dat resident in HBM."""
    import ops_amd
    from ops_amd.apps.cloverleaf3d.decks import GOLD_KE_BM, bm_deck
    L = ops_amd.app("cloverleaf3d")
    out = {}
    decks = [d for d in (("deck_96", 96, 87), ("large", args.c3_cells, args.c3_steps))
             if d[0] in args.c3_decks.split(",")]
    for tag, n, steps in decks:
        with tempfile.TemporaryDirectory() as t:
            deck = Path(t) / "clover.in"
            deck.write_text(bm_deck(n, steps))
            argc, argv = ops_amd.argv()
            log(f"[bench] cloverleaf3d {n}^3: warm-up")
            assert L.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
            # 5 candidate shapes x 2 launches; some loops run every other step
            assert L.cloverleaf3d_step(24) == 0
            assert L.cloverleaf3d_finalize() == 0
            graph = not args.no_graph
            assert L.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
            if graph:
                L.cloverleaf3d_set_graph(1)
                assert L.cloverleaf3d_capture() == 0
            R.ops_hip_device_sync()
            if not graph:
                R.ops_hip_reset_stats()
                R.ops_hip_set_timing(1)
            log(f"[bench] cloverleaf3d {n}^3: timed {steps} steps")
            t0 = time.perf_counter()
            rc = L.cloverleaf3d_step(steps)
            R.ops_hip_device_sync()
            t1 = time.perf_counter()
            R.ops_hip_set_timing(0)
            L.cloverleaf3d_set_graph(0)
            assert rc == 0
            s = (ctypes.c_double * 5)()
            assert L.cloverleaf3d_summary(s) == 0
            assert L.cloverleaf3d_finalize() == 0
            eager_s = None
            if graph:
                assert L.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
                R.ops_hip_device_sync()
                R.ops_hip_reset_stats()
                R.ops_hip_set_timing(1)
                e0 = time.perf_counter()
                assert L.cloverleaf3d_step(steps) == 0
                R.ops_hip_device_sync()
                eager_s = time.perf_counter() - e0
                R.ops_hip_set_timing(0)
                assert L.cloverleaf3d_finalize() == 0
            stats = ops_amd.kernel_stats()
        secs = t1 - t0
        e = {"workload": f"cloverleaf_3d clover.in {n}^3 cells, {steps} hydro steps",
             "s": round(secs, 4), "ms_per_step": round(secs * 1e3 / steps, 3),
             "cell_updates_per_s": n ** 3 * steps / secs,
             "step_bandwidth": whole_step_bw(stats, secs),
             "dominant_kernel": roofline(stats)[1], "ke": s[3],
             "launches_per_step": round(sum(v["calls"] for v in stats.values()) / steps, 1),
             "top_kernels": top_kernels(stats, 8), "captured_steps": graph}
        if eager_s is not None:
            e["eager_ms_per_step_with_kernel_timing"] = round(eager_s * 1e3 / steps, 3)
        if n == 96 and steps == 87:
            e["ke_golden"] = GOLD_KE_BM
            e["ke_pct_diff"] = abs(100.0 * (s[3] / GOLD_KE_BM) - 100.0)
            e["ke_pass_0.001pct"] = e["ke_pct_diff"] < 0.001
        out[tag] = e
    if not args.no_cpu:
        cpu = host_cpu()
        lib = cpu_lib("cloverleaf3d", cpu["threads"])
        steps = args.c3_cpu_steps
        with tempfile.TemporaryDirectory() as t:
            deck = Path(t) / "clover.in"
            deck.write_text(bm_deck(96, 87))
            argc, argv = ops_amd.argv()
            assert lib.cloverleaf3d_init(argc, argv, str(deck).encode()) == 0
            dt = median_time(lambda: lib.cloverleaf3d_step(steps))
            assert lib.cloverleaf3d_finalize() == 0
        out["cpu_port"] = {"value": 96 ** 3 * steps / dt, "unit": "cell-updates/s",
                           "cores": cpu["threads"], "kind": "port", "cpu": cpu,
                           "sample": f"CloverLeaf 3D 96^3, clover.in, median of 3 runs of {steps} "
                                     f"steps after a warm-up run on the oracle host backend "
                                     f"(OpenMP, {cpu['threads']} threads): {dt:.2f} s"}
    return out


def clover_cpu(args, cpu):
    """Oracle host backend, same application and deck, bounded sample of steps: one warm-up
    step, then the median of three runs of --cpu-steps steps."""
    import ops_amd
    lib = cpu_lib("cloverleaf", cpu["threads"])
    n = args.cells
    steps = args.cpu_steps
    with tempfile.TemporaryDirectory() as t:
        deck = Path(t) / "clover.in"
        deck.write_text(BM16_DECK.format(n=n))
        argc, argv = ops_amd.argv()
        assert lib.cloverleaf_init(argc, argv, str(deck).encode()) == 0
        assert lib.cloverleaf_step(1) == 0  # warm-up step (page faults, first touch)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            assert lib.cloverleaf_step(steps) == 0
            ts.append(time.perf_counter() - t0)
        assert lib.cloverleaf_finalize() == 0
    dt = sorted(ts)[1]
    return {"value": n * n * steps / dt, "unit": "cell-updates/s", "cores": cpu["threads"],
            "kind": "port", "cpu": cpu,
            "sample": f"CloverLeaf {n}x{n}, clover_bm16_short.in, after one warm-up step the median "
                      f"of 3 runs of {steps} steps on the oracle host backend (OpenMP, "
                      f"{cpu['threads']} threads): {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=87)
    ap.add_argument("--warmup", type=int, default=24,
                    help="untimed steps first (the autotuner tries its candidate shapes there)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the runtime's default workgroup shapes (no OPS_AUTOTUNE)")
    ap.add_argument("--workload", default="cloverleaf", choices=["cloverleaf"])
    ap.add_argument("--cells", type=int, default=3840)
    ap.add_argument("--chain", type=int, default=0,
                    help="CloverLeaf: run the loops between halo exchanges as loop chains of "
                         "this many bands (0: one launch per loop)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="CloverLeaf 2D steps per timed CPU-baseline run (median of 3)")
    ap.add_argument("--c3-cpu-steps", type=int, default=5)
    ap.add_argument("--no-laplace", action="store_true")
    ap.add_argument("--lap-cells", type=int, default=1024)
    ap.add_argument("--lap-iters", type=int, default=1000)
    ap.add_argument("--lap-cpu-iters", type=int, default=200)
    ap.add_argument("--tea-solvers", type=lambda v: [int(x) for x in v.split(",")],
                    default=[0, 1, 3, 4],
                    help="TeaLeaf solvers to time (0 plain CG, 1 device CG, 2 FP32 CG + FP64 "
                         "refinement, 3 FP64 PCG + FP32 multigrid, 4 FP64 PCG + FP64 multigrid)")
    ap.add_argument("--tea-cpu-iters", type=int, default=30)
    ap.add_argument("--adi-cpu-blocks", type=int, default=6)
    ap.add_argument("--no-tealeaf", action="store_true")
    ap.add_argument("--no-adi", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--unet-pairs", type=int, default=2000)
    ap.add_argument("--unet-synthetic", type=int, default=0,
                    help="also time U-Net training alone on this many random 1024^2 samples")
    ap.add_argument("--unet-epochs", type=int, default=4)
    ap.add_argument("--unet-steps-per-deck", type=int, default=10)
    ap.add_argument("--adi-steps", type=int, default=100)
    ap.add_argument("--adi-eager", action="store_true",
                    help="time eager launches instead of the captured step (for profilers)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time CloverLeaf steps as eager launches (default: replayed hipGraphs)")
    ap.add_argument("--no-clover3d", action="store_true")
    ap.add_argument("--c3-cells", type=int, default=320)
    ap.add_argument("--c3-steps", type=int, default=20)
    ap.add_argument("--c3-decks", default="deck_96,large",
                    help="CloverLeaf 3D meshes to run: deck_96 (the reference's deck), large")
    ap.add_argument("--tea-cells", type=int, default=4096)
    ap.add_argument("--tea-steps", type=int, default=10)
    ap.add_argument("--ingest-pmc", metavar="DIR", default=None,
                    help="read rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE CSV output under DIR "
                         "and write profiles/pmc_traffic.json, then exit")
    ap.add_argument("--ingest-match", metavar="REGEX", default=None,
                    help="with --ingest-pmc: take only the kernels whose full name matches REGEX")
    args = ap.parse_args()
    if args.ingest_pmc:
        ingest_pmc(Path(args.ingest_pmc), args.ingest_match)
        return
    if not args.no_autotune:
        # the runtime measures candidate workgroup shapes (and tridiagonal tile heights) on
        # the warm-up launches and keeps the fastest for the timed ones
        os.environ["OPS_AUTOTUNE"] = "1"
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: runs on one GPU only; the mesh is not decomposed across GPUs, "
              "so --gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    import ops_amd
    R = ops_amd.runtime()
    secs, stats, checks = clover_gpu(args, R)
    n = args.cells
    value = n * n * args.steps / secs
    kname, roof = roofline(stats)
    if checks.get("hbm_copy_GBps", 0) > 0:
        roof["measured_copy_GBps"] = checks["hbm_copy_GBps"]
        roof["frac_of_measured_copy"] = round(roof["achieved"] / checks["hbm_copy_GBps"], 4)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": secs * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cloverleaf_2d clover_bm16_short.in {n}x{n} cells, "
                               f"{args.steps} hydro steps", "cells": n * n,
                   "halo_depth": 2, "parallelism": "single GPU",
                   "autotune": not args.no_autotune},
        "roofline": roof,
        "step_bandwidth": whole_step_bw(stats, secs),
        "top_kernels": top_kernels(stats, 10),
        "checks": checks,
    }
    cpu = None if args.no_cpu else host_cpu()
    if not args.no_laplace:
        log("[bench] laplace2d")
        out["laplace"] = laplace_gpu(args)
        if cpu:
            out["laplace"]["cpu_baseline"] = laplace_cpu(args, cpu)
    if not args.no_tealeaf:
        out["tealeaf"] = tealeaf_gpu(args)
        if cpu:
            log("[bench] tealeaf cpu sample")
            out["tealeaf"]["cpu_baseline"] = tc = tealeaf_cpu(args, cpu)
            plain = out["tealeaf"]["iterations"].get("fp64_cg_plain")
            if plain:
                tc["time_to_solution_s_extrapolated"] = round(tc["s_per_cg_iteration"] * plain, 2)
    if not args.no_learned:
        out["tealeaf_learned"] = tealeaf_learned_gpu(args)
    if args.unet_synthetic > 0:
        out["unet_training"] = unet_training_gpu(args)
    if not args.no_adi:
        out["adi"] = adi_gpu(args)
        if cpu:
            log("[bench] adi cpu sample")
            out["adi"]["cpu_baseline"] = adi_cpu(args, cpu)
    if not args.no_clover3d:
        out["cloverleaf_3d"] = clover3d_gpu(args, R)
    if cpu:
        log(f"[bench] cpu baseline: {cpu['threads']} threads, 3 x {args.cpu_steps} steps")
        out["cpu_baseline"] = clover_cpu(args, cpu)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

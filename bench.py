#!/usr/bin/env python3
"""Benchmark: Riemann3 128^3 single solve on one MI355X vs the reference CPU Fast Marching.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one complete solve of the Riemann3 workload (128^3 grid, h = 1/128, one seed at the
centre, SPD metric R diag(1, a^2, b^2) R^T with a, b in [1, 20] from smooth seeded fields):
FP64 Selling stencils from the HBM-resident metric, block fast-iterative FP32 update to the
fixed point, and the row-major FP64 'values' output, all on the device.  Inputs are resident in
HBM before the timed region; the timed region is bracketed by torch.cuda.synchronize().

Printed JSON fields beyond the driver's contract:
  roofline      update kernel (k_fim): algorithmic bytes per launch / mean launch duration
                (hipEvent pairs around every launch on the plan's stream, in a separate
                instrumented solve), against the 8 TB/s HBM peak.  The unit of work is the
                point-update (one local solve at one point in one tile sweep) at 50 B each;
                achieved_per_visit counts each tile visit once instead (its data is read from
                HBM once and swept in LDS/registers), and solve_effective_gbs divides the
                stencil + value bytes of the whole grid by the solve time.  traffic = HBM bytes
                per launch from two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) run by this
                script over a short run of itself, each scaled by what it reads for a plain
                1 GiB device copy (the committed profiles/fim_riemann3_128_traffic.json only
                when rocprofv3 is absent); copy_ceiling_gbs = a timed 1 GiB device copy,
                (read + write bytes) / time, the practical ceiling beside the 8 TB/s peak.
  serving       the continuous-batching server on a stream of mixed Riemann2 / Riemann3
                requests (70 % 2-D, side 64-1024 log-uniform; 30 % 3-D, side 32-160; 1-16
                seeds; anisotropy 1-50): sustained solves/s with 128 requests kept in flight
                (closed loop, window excluding the first and last 128 completions), the rate
                of a burst with every request submitted at once (makespan), p50 / p99
                latency at an offered Poisson load of 5x the one-at-a-time rate, and the
                one-at-a-time plain GPU solves/s of the same requests; served_sample_check:
                served results of the largest 2-D request and three smaller ones against the
                oracle's FP64 solve, every point.
  cpu_baseline  the reference's FileHFM_Riemann3, compiled from its sources into oracle/_ref,
                timed on this host (median of 3 full solves, 1 core; StencilCPUTime + FMCPUTime),
                or the oracle's C++ restatement when the reference binary is absent; with the
                host's CPU model.
  accuracy      max relative error of the FP32 values vs the reference FP64 values (floored at
                the smallest nonzero reference value), from the CPU-baseline run.
  configs       the other workloads on the GPU (ms per solve from HBM-resident inputs; geodesics
                and adjoint timed separately) beside the reference executable on the same
                inputs (StencilCPUTime + FMCPUTime, 1 core) and its accuracy.
  io_path       the drop-in IO path (HFMIO set + hfm_run + values read back) on the same
                workload, host buffers in and out: never the headline value.
  residual      final FP64 residual check of the last timed solve (hfm_plan_residual): the
                largest relative deviation between each reached non-seed value and the FP64
                root of its local scheme, and the finite / infinite mismatches.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/solve Riemann3 128^3 vs HFM CPU; served solves/s + p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per point-update of the FP32 Riemann3 update: u read 4 + u write 4 +
# 6 FP32 weights 24 + 6 x 3 int8 offsets 18 (neighbour reads are cache hits).
BYTES_PER_POINT_UPDATE = 50
# Whole-grid bytes of one solve: metric in 48 + stencil out 42 (stencil kernel), stencil in 42
# + value out 4 (update), values out 8 (export).
BYTES_PER_POINT_SOLVE = 48 + 42 + 42 + 4 + 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serving-requests", type=int, default=300)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (Riemann2 256^2 FP64, ReedsSheppForward2 256x256x64 with 32 "
                         "geodesics, Elastica2 192x192x64 with its adjoint)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device-to-device copy (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="do not measure the update kernel's HBM traffic with rocprofv3 (use the committed "
                         "profile's figure)")
    return ap.parse_args()


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), as lscpu reports it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, runs=3):
    """Time full CPU solves: the reference binary if built, else the oracle restatement; the
    median of `runs` runs (StencilCPUTime + FMCPUTime, clock() CPU time, 1 core)."""
    import oracle
    cores = 1
    use_ref = oracle.ref_available(params["model"])
    times, fms, walls = [], [], []
    values = None
    for _ in range(runs):
        t = time.time()
        out = oracle.run_reference(params) if use_ref else oracle.solve(params)
        walls.append(time.time() - t)
        if use_ref and out.get("retcode", 1) != 0:
            raise RuntimeError("reference run failed: " + out.get("log", "")[-500:])
        times.append(out["StencilCPUTime"] + out["FMCPUTime"])
        fms.append(out["FMCPUTime"])
        values = out["values"]
    i = int(np.argsort(times)[len(times) // 2])
    return {
        "value": round(times[i] * 1e3, 1), "unit": "ms/solve", "cores": cores,
        "kind": "reference" if use_ref else "port",
        "sample": f"median of {runs} full Riemann3 {params['dims'][0]:.0f}^3 solves (StencilCPUTime + FMCPUTime, "
                  f"clock() CPU time, 1 core), wall {sum(walls):.1f} s in all",
        "runs_ms": [round(t * 1e3, 1) for t in times],
        "fm_ms": round(fms[i] * 1e3, 1), "nproc": os.cpu_count(), "cpu_model": cpu_model(),
    }, values


def measure_traffic(timeout_s=180):
    """HBM bytes per launch of the update kernel k_fim, measured now: two rocprofv3 --pmc passes
    (FETCH_SIZE, then WRITE_SIZE) over a short run of this benchmark (2 timed solves and a plain
    1 GiB device copy), each counter scaled by what it reads for the copy.  Returns (result, note)."""
    import glob
    import shutil
    import subprocess
    import tempfile
    from hamiltonfastmarching_amd import profiling
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    child = [sys.executable, "-u", os.path.abspath(__file__), "--steps", "2", "--warmup", "1", "--no-cpu",
             "--no-serving", "--no-configs", "--copy-calibration", "--no-traffic"]
    with tempfile.TemporaryDirectory(prefix="hfm_pmc_") as td:
        dbs = []
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(td, ctr)
            r = subprocess.run(["timeout", "-s", "KILL", str(timeout_s), exe, "--pmc", ctr, "-d", d, "-o", "run", "--",
                                *child], capture_output=True, text=True)
            found = glob.glob(os.path.join(d, "**", "run_results.db"), recursive=True)
            if not found:
                return None, f"rocprofv3 --pmc {ctr}: rc {r.returncode}, no results ({r.stderr[-300:]})"
            dbs.append(found[0])
        res = profiling.counters(dbs[0], dbs[1], "k_fim", float(1 << 30))
    return res, "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in this run, scaled by a 1 GiB device copy"


def io_path_leg(params, reps=3):
    """The drop-in IO path on the same Riemann3 workload (FP32 update): the input keys copied into
    an HFMIO object (set), hfm_run (run: host buffers in, upload, stencils, solve, exports back to
    the host), and the values read back (get); medians of `reps` calls.  Never the headline value:
    it includes host-memory copies and PCIe transfers that the device-resident path does not."""
    import hamiltonfastmarching_amd as hfm
    q = dict(params, precision=32.0)
    hfm.run(q)
    rows = []
    for _ in range(reps):
        t0 = time.perf_counter()
        io = hfm.HFMIO(default_ordering="")
        for k, v in q.items():
            io[k] = v
        t1 = time.perf_counter()
        io.run()
        t2 = time.perf_counter()
        _ = io["values"]
        t3 = time.perf_counter()
        rows.append(((t3 - t0) * 1e3, (t1 - t0) * 1e3, (t2 - t1) * 1e3, (t3 - t2) * 1e3))
    a = np.median(np.array(rows), axis=0)
    return {"io_path_ms": round(float(a[0]), 2), "set_ms": round(float(a[1]), 2), "run_ms": round(float(a[2]), 2),
            "get_ms": round(float(a[3]), 2),
            "definition": "HFMIO set of the input keys (a copy into the IO store, as the reference's interface "
                          "does) + hfm_run + values read back, median of 3, FP32 update"}


def copy_ceiling_gbs(nbytes=1 << 30, reps=10):
    """Practical HBM ceiling: a plain device-to-device copy of nbytes, (read + write bytes) / time."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def serving_leg(count, with_cpu=False):
    """Continuous-batching server vs one-at-a-time plain GPU solves on the same request stream."""
    import torch
    from hamiltonfastmarching_amd import serving
    reqs = serving.make_requests(count, seed=11)
    serving.run_sequential(reqs[:4])
    seq = serving.run_sequential(reqs)
    srv = serving.Server(hbm_budget_bytes=64e9, max_resident=256)
    try:
        serving.run_load(srv, reqs[:16], qps=0)
        check = served_sample_check(srv, reqs)
        burst = serving.run_load(srv, reqs, qps=0)
        closed = serving.run_closed_loop(srv, reqs, concurrency=128, rounds=3)
        offered = 5.0 * seq["solves_per_s"]
        load = serving.run_load(srv, reqs, qps=offered)
        stats = srv.stats()
    finally:
        srv.close()
    torch.cuda.synchronize()
    n2 = sum(r["model"] == "Riemann2" for r in reqs)
    sustained = closed["sustained_solves_per_s"]
    cpu = cpu_serving(reqs) if with_cpu else None
    return {
        "requests": count, "mix": {"Riemann2": n2, "Riemann3": count - n2},
        "solves_per_s": round(sustained, 1),
        "solves_per_s_definition": "sustained completion rate with 128 requests kept in flight (closed "
                                   "loop over 3x the request stream; completions 128 .. 3n-128)",
        "closed_loop_latency": {"p50_ms": round(closed["p50_ms"], 3), "p99_ms": round(closed["p99_ms"], 3)},
        "sequential_solves_per_s": round(seq["solves_per_s"], 1),
        "speedup_vs_sequential": round(sustained / seq["solves_per_s"], 2),
        "burst": {"solves_per_s": round(burst["solves_per_s"], 1),
                  "speedup_vs_sequential": round(burst["solves_per_s"] / seq["solves_per_s"], 2),
                  **{k: round(burst[k], 3) for k in ("p50_ms", "p99_ms", "makespan_s")}},
        "at_5x_sequential_rate": {"qps_offered": round(offered, 1), "solves_per_s": round(load["solves_per_s"], 1),
                                  "p50_ms": round(load["p50_ms"], 3), "p99_ms": round(load["p99_ms"], 3)},
        "sequential_latency": {"p50_ms": round(seq["p50_ms"], 3), "p99_ms": round(seq["p99_ms"], 3)},
        "max_resident_seen": stats["max_resident_seen"],
        "served_sample_check": check,
        "cpu_aggregate": cpu,
        "speedup_vs_cpu_aggregate": round(sustained / cpu["solves_per_s"], 1) if cpu and cpu["solves_per_s"] else None,
    }


def served_sample_check(srv, reqs, count=4, max_points=300_000):
    """Results of the server checked against the oracle's FP64 Fast Marching (all points of each
    request): the largest 2-D request of the stream and `count` - 1 others drawn at random among
    those of at most max_points points (the oracle is a single-core CPU solve)."""
    import oracle
    from hamiltonfastmarching_amd import serving
    rng = np.random.Generator(np.random.PCG64(5))
    two = [i for i, r in enumerate(reqs) if r["model"] == "Riemann2"]
    pick = [max(two, key=lambda i: int(np.prod(reqs[i]["dims"])))] if two else []
    small = [i for i, r in enumerate(reqs) if int(np.prod(r["dims"])) <= max_points and i not in pick]
    pick += [int(i) for i in rng.choice(small, size=min(count - len(pick), len(small)), replace=False)]
    sub = [reqs[i] for i in pick]
    res = serving.run_load(srv, sub, qps=0, collect_values=True)
    worst = 0.0
    cases = []
    for r, u in zip(sub, res["values"]):
        ref = oracle.solve({"model": r["model"], "dims": list(r["dims"]), "gridScale": r["gridScale"],
                            "metric": r["metric"].cpu().numpy(), "seeds": r["seeds"]})["values"]
        same = bool(np.array_equal(np.isfinite(u), np.isfinite(ref)))
        err = floor_rel(u, ref) if same else float("inf")
        worst = max(worst, err)
        cases.append({"model": r["model"], "dims": list(r["dims"]), "kappa": round(r["kappa"], 1),
                      "seeds": len(r["seeds"]), "max_rel_err": err})
    return {"max_rel_err_vs_oracle": worst, "tolerance": 1e-5, "cases": cases}


def cpu_serving(reqs, budget_s=15.0, max_workers=16):
    """The CPU side of the serving comparison: the same request stream solved by independent
    reference (or oracle) runs on the host cores, one process per core, for about budget_s
    seconds (workers stop taking requests after that).  Returns the aggregate solves/s."""
    import threading
    import oracle
    workers = max(1, min(max_workers, os.cpu_count() or 1))
    use_ref = all(oracle.ref_available(m) for m in ("Riemann2", "Riemann3"))
    host = [{"model": r["model"], "dims": np.array(r["dims"], float), "gridScale": r["gridScale"],
             "metric": r["metric"].cpu().numpy(), "seeds": np.asarray(r["seeds"])} for r in reqs]
    lock = threading.Lock()
    state = {"next": 0, "done": 0, "failed": 0}
    t0 = time.time()

    def work():
        while True:
            with lock:
                if state["next"] >= len(host) or time.time() - t0 > budget_s:
                    return
                i = state["next"]
                state["next"] += 1
            if use_ref:
                ok = oracle.run_reference(host[i]).get("retcode", 1) == 0
            else:
                oracle.solve(host[i])
                ok = True
            with lock:
                state["done" if ok else "failed"] += 1

    threads = [threading.Thread(target=work) for _ in range(workers)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.time() - t0
    return {"solves_per_s": round(state["done"] / wall, 2), "workers": workers, "requests": state["done"],
            "failed": state["failed"], "kind": "reference" if use_ref else "port",
            "sample": f"first {state['done']} requests of the serving stream, {workers} concurrent single-core "
                      f"solves, {wall:.1f} s wall"}


def _time_gpu(fn, steps, warmup=1):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def configs_leg(with_cpu):
    """The other workloads, each timed on the GPU from HBM-resident inputs and, with with_cpu,
    beside the reference executable on the same inputs (StencilCPUTime + FMCPUTime, 1 core):
      riemann2_256_fp64    Riemann2 256^2, h = 1/256, FP64 update, centre seed
      rsf2_256x256x64      ReedsSheppForward2, xi = 0.1, FP32, 4 seeds; then 32 geodesics by the
                           discrete solver and by the ODE solver
      elastica2_192x192x64 Elastica2, xi = 0.2, FP32, 1 seed; then the adjoint of the sum of the
                           arrival times at 32 targets (costSensitivity)
      riemann2_256_order2_serial  Riemann2 256^2 at order 2 (the serial acceptance loop), FP64,
                           beside the reference's FMCPUTime"""
    import torch
    import oracle
    import hamiltonfastmarching_amd as hfm
    from hamiltonfastmarching_amd import synthetic as syn

    def cpu(params, keys=("StencilCPUTime", "FMCPUTime")):
        if not with_cpu or not oracle.ref_available(params["model"]):
            return None, None
        t = time.time()
        out = oracle.run_reference(params)
        wall = time.time() - t
        if out.get("retcode", 1) != 0:
            return None, None
        return {"ms": round(sum(float(out[k]) for k in keys) * 1e3, 1), "keys": "+".join(keys),
                "wall_ms": round(wall * 1e3, 1), "cores": 1}, out

    res = {}
    # Riemann2 256^2 FP64
    p = syn.riemann2(256, seed=0)
    plan = hfm.Plan("Riemann2", (256, 256), p["gridScale"], precision=64)
    m = torch.from_numpy(p["metric"]).cuda()
    out = torch.empty((256, 256), dtype=torch.float64, device="cuda")
    ms = _time_gpu(lambda: (plan.set_metric(m), plan.solve(p["seeds"]), plan.values(out)), 10)
    c, ref = cpu(p)
    r = {"gpu_ms": round(ms, 3), "cpu": c}
    if ref is not None:
        r["max_rel_err_vs_reference"] = floor_rel(out.cpu().numpy(), np.asarray(ref["values"]).reshape(256, 256))
        r["speedup"] = round(c["ms"] / ms, 1)
    res["riemann2_256_fp64"] = r
    del plan
    # ReedsSheppForward2 256 x 256 x 64, FP32, geodesics from 32 tips
    full = syn.rsf2(256, 64, seed=0)
    tips = full.pop("tips")
    dims = (256, 256, 64)
    plan = hfm.Plan("ReedsSheppForward2", dims, full["gridScale"], origin=full["origin"], precision=32,
                    xi=float(full["xi"]), eps=float(full["eps"]))
    cost = torch.from_numpy(full["cost"]).cuda()
    out = torch.empty(dims, dtype=torch.float64, device="cuda")
    ms = _time_gpu(lambda: (plan.set_cost(cost), plan.solve(full["seeds"]), plan.values(out)), 3)
    gms = _time_gpu(lambda: plan.geodesics(tips), 2)
    c, ref = cpu(dict(full, tips=tips), keys=("StencilCPUTime", "FMCPUTime"))
    r = {"gpu_ms": round(ms, 3), "geodesics_gpu_ms": round(gms, 3), "cpu": c}
    if ref is not None:
        r["geodesics_cpu_ms"] = round(float(ref["GeodesicCPUTime"]) * 1e3, 1)
        r["max_rel_err_vs_reference"] = floor_rel(out.cpu().numpy(), np.asarray(ref["values"]).reshape(dims))
        r["speedup"] = round(c["ms"] / ms, 1)
    # the same 32 tips through the ODE geodesic solver (geodesicSolver = "ODE")
    oms = _time_gpu(lambda: plan.geodesics_ode(tips), 2)
    r["geodesics_ode_gpu_ms"] = round(oms, 3)
    if c is not None:
        _, oref = cpu(dict(full, tips=tips, geodesicSolver="ODE"))
        if oref is not None:
            r["geodesics_ode_cpu_ms"] = round(float(oref["GeodesicCPUTime"]) * 1e3, 1)
            _, olens, _ = plan.geodesics_ode(tips)
            r["geodesics_ode_lengths_equal"] = float(np.mean(olens == np.asarray(oref["geodesicLengths"])))
    res["rsf2_256x256x64"] = r
    del plan
    # Elastica2 192 x 192 x 64, FP32, adjoint for 32 targets
    full = syn.elastica2(192, 64, seed=0)
    targets = full["inspectSensitivity"]
    dims = (192, 192, 64)
    plan = hfm.Plan("Elastica2", dims, full["gridScale"], origin=full["origin"], precision=32,
                    xi=float(full["xi"]), eps=float(full["eps"]))
    cost = torch.from_numpy(full["cost"]).cuda()
    ms = _time_gpu(lambda: (plan.set_cost(cost), plan.solve(full["seeds"])), 3)
    ams = _time_gpu(lambda: plan.adjoint(targets), 3)
    u = plan.values().cpu().numpy()
    sens = plan.adjoint(targets)[0].cpu().numpy()
    c, ref = cpu(full)
    r = {"gpu_ms": round(ms, 3), "adjoint_gpu_ms": round(ams, 3), "cpu": c}
    if ref is not None:
        r["speedup_solve"] = round(c["ms"] / ms, 1)
        r["cpu_wall_ms_with_backward"] = c["wall_ms"]
        r["max_rel_err_vs_reference"] = floor_rel(u, np.asarray(ref["values"]).reshape(dims))
        sref = np.asarray(ref["costSensitivity_0"]).reshape(dims)
        r["adjoint_max_err_vs_reference"] = float(np.abs(sens - sref).max() / np.abs(sref).max())
        r["adjoint_err_definition"] = "max |costSensitivity - reference| / max |reference| (FP32 solve and adjoint)"
    res["elastica2_192x192x64"] = r
    del plan
    # Riemann2 256^2 at order 2: the serial acceptance loop, bit-equal to the reference
    p = dict(syn.riemann2(256, seed=0), order=2.0)
    plan = hfm.Plan("Riemann2", (256, 256), p["gridScale"], precision=64)
    m = torch.from_numpy(p["metric"]).cuda()
    plan.set_metric(m)
    plan.set_order(2)
    out = torch.empty((256, 256), dtype=torch.float64, device="cuda")
    ms = _time_gpu(lambda: (plan.solve(p["seeds"]), plan.values(out)), 2)
    c, ref = cpu(p, keys=("FMCPUTime",))
    r = {"gpu_solve_ms": round(ms, 3), "cpu": c, "note": "one wave runs the reference's heap loop"}
    if ref is not None:
        r["bit_equal_to_reference"] = bool(np.array_equal(out.cpu().numpy(),
                                                          np.asarray(ref["values"]).reshape(256, 256)))
        r["gpu_over_cpu_time"] = round(ms / c["ms"], 1)
    res["riemann2_256_order2_serial"] = r
    del plan
    torch.cuda.synchronize()
    return res


def floor_rel(u, ref):
    fin = np.isfinite(ref)
    pos = ref[fin & (ref > 0)]
    floor = pos.min()
    m = fin & (ref >= floor)
    return float(np.max(np.abs(u[m] - ref[m]) / ref[m]))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "a multi-GPU layout is described but not built.", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    import hamiltonfastmarching_amd as hfm
    from hamiltonfastmarching_amd import synthetic as syn

    n = args.n
    params = syn.riemann3(n, seed=0)
    metric = torch.from_numpy(params["metric"]).cuda()
    seeds = params["seeds"]
    plan = hfm.Plan("Riemann3", (n, n, n), params["gridScale"], precision=32)
    out = torch.empty((n, n, n), dtype=torch.float64, device="cuda")

    def step():
        plan.set_metric(metric)
        plan.solve(seeds)
        plan.values(out)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / max(args.steps, 1)
    st = plan.stats()
    # final FP64 residual check of the last timed solve (outside the timed region)
    res_dev, res_bad = plan.residual()

    # Instrumented solve: per-launch hipEvent timing of the update kernel.
    plan.set_timing(True)
    step()
    torch.cuda.synchronize()
    sti = plan.stats()
    plan.set_timing(False)
    launches = max(sti["fim_launches"], 1)
    mean_launch_ms = sti["fim_ms"] / launches

    def gbs(total_bytes):
        return (total_bytes / launches) / (mean_launch_ms * 1e-3) / 1e9 if mean_launch_ms > 0 else 0.0

    achieved = gbs(sti["sweeps"] * 64 * BYTES_PER_POINT_UPDATE)
    achieved_visit = gbs(sti["tile_visits"] * 64 * BYTES_PER_POINT_UPDATE)
    traffic, traffic_source = None, None
    if n == 128 and not args.no_traffic:
        tr, note = measure_traffic()
        if tr is not None and tr.get("hbm_bytes_per_launch") == tr.get("hbm_bytes_per_launch"):
            traffic, traffic_source = tr["hbm_bytes_per_launch"], note
        else:
            traffic_source = "not measured: " + note
    if traffic is None and n == 128:
        tpath = os.path.join(ROOT, "profiles", "fim_riemann3_128_traffic.json")
        if os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
            traffic_source = (traffic_source + "; " if traffic_source else "") + \
                "profiles/fim_riemann3_128_traffic.json (committed profile)"
    ceiling = copy_ceiling_gbs()
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_source,
        "traffic_gbs": round(traffic / (mean_launch_ms * 1e-3) / 1e9, 1) if traffic and mean_launch_ms > 0 else None,
        "frac_traffic": round(traffic / (mean_launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if traffic and mean_launch_ms > 0 else None,
        "copy_ceiling_gbs": round(ceiling, 1),
        "kernel": "k_fim<Riemann3,float>", "launches": sti["fim_launches"],
        "mean_launch_us": round(mean_launch_ms * 1e3, 2), "unit_of_work": "point-update",
        "bytes_per_point_update": BYTES_PER_POINT_UPDATE, "point_updates": sti["sweeps"] * 64,
        "tile_visits": sti["tile_visits"], "achieved_per_visit": round(achieved_visit, 1),
        "frac_per_visit": round(achieved_visit / HBM_PEAK_GBS, 4),
        "solve_effective_gbs": round(n ** 3 * BYTES_PER_POINT_SOLVE / (ms * 1e-3) / 1e9, 1),
    }
    result = {
        "metric": METRIC, "value": round(ms, 3), "unit": "ms/solve", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded smooth SPD field, anisotropy up to 20)",
        "config": {"workload": f"Riemann3 {n}^3 single solve, h=1/{n}, centre seed, FP32 update, FP64 Selling",
                   "grid": [n, n, n], "seeds": 1},
        "solves_per_s": round(1e3 / ms, 2),
        "solve_breakdown_ms": {"stencil_and_deps": round(st["stencil_ms"], 3), "fim": round(st["solve_ms"], 3)},
        "fim_iterations": st["iterations"], "tile_visits": st["tile_visits"], "sweeps": st["sweeps"],
        "roofline": roofline,
        "residual": {"max_rel_dev": res_dev, "mismatches": res_bad,
                     "definition": "FP64 recompute of the local scheme at every reached non-seed point: "
                                   "max |Lambda(u) - u| / |u|"},
    }
    if n == 128 and not args.no_configs:
        result["io_path"] = io_path_leg(params)
    if not args.no_serving:
        result["serving"] = serving_leg(args.serving_requests, with_cpu=not args.no_cpu)
    if not args.no_configs:
        result["configs"] = configs_leg(with_cpu=not args.no_cpu)
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        result["copy_calibration_bytes"] = src.numel() * 4
        del src, dst
    if not args.no_cpu:
        cb, ref_values = cpu_baseline(params)
        result["cpu_baseline"] = cb
        u = out.cpu().numpy()
        result["accuracy"] = {"max_rel_err_vs_fp64_reference": floor_rel(u, ref_values),
                              "tolerance": 1e-5}
        result["speedup_vs_cpu"] = round(cb["value"] / ms, 1)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()

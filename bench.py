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
                per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy
                (profiles/fim_riemann3_128_traffic.json).
  serving       the continuous-batching server on a stream of mixed Riemann2 / Riemann3
                requests (70 % 2-D, side 64-1024 log-uniform; 30 % 3-D, side 32-160; 1-16
                seeds; anisotropy 1-50): sustained solves/s with 128 requests kept in flight
                (closed loop, window excluding the first and last 128 completions), the rate
                of a burst with every request submitted at once (makespan), p50 / p99
                latency at an offered Poisson load of 5x the one-at-a-time rate, and the
                one-at-a-time plain GPU solves/s of the same requests.
  cpu_baseline  the reference's FileHFM_Riemann3, compiled from its sources into oracle/_ref,
                timed on this host (one full solve, 1 core; StencilCPUTime + FMCPUTime), or the
                oracle's C++ restatement when the reference binary is absent.
  accuracy      max relative error of the FP32 values vs the reference FP64 values (floored at
                the smallest nonzero reference value), from the CPU-baseline run.
  configs       the other workloads on the GPU (ms per solve from HBM-resident inputs; geodesics
                and adjoint timed separately) beside the reference executable on the same
                inputs (StencilCPUTime + FMCPUTime, 1 core) and its accuracy.
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
    return ap.parse_args()


def cpu_baseline(params):
    """Time one full CPU solve: the reference binary if built, else the oracle restatement."""
    import oracle
    cores = 1
    if oracle.ref_available(params["model"]):
        t = time.time()
        out = oracle.run_reference(params)
        wall = time.time() - t
        if out.get("retcode", 1) != 0:
            raise RuntimeError("reference run failed: " + out.get("log", "")[-500:])
        cpu_s = out["StencilCPUTime"] + out["FMCPUTime"]
        kind = "reference"
        values = out["values"]
        fm_s = out["FMCPUTime"]
    else:
        t = time.time()
        out = oracle.solve(params)
        wall = time.time() - t
        cpu_s = out["StencilCPUTime"] + out["FMCPUTime"]
        kind = "port"
        values = out["values"]
        fm_s = out["FMCPUTime"]
    return {
        "value": round(cpu_s * 1e3, 1), "unit": "ms/solve", "cores": cores, "kind": kind,
        "sample": f"1 full Riemann3 {params['dims'][0]:.0f}^3 solve (StencilCPUTime + FMCPUTime, "
                  f"clock() CPU time), wall {wall:.1f} s",
        "fm_ms": round(fm_s * 1e3, 1), "nproc": os.cpu_count(),
    }, values


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
        "cpu_aggregate": cpu,
        "speedup_vs_cpu_aggregate": round(sustained / cpu["solves_per_s"], 1) if cpu and cpu["solves_per_s"] else None,
    }


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
      rsf2_256x256x64      ReedsSheppForward2, xi = 0.1, FP32, 4 seeds; then 32 geodesics
      elastica2_192x192x64 Elastica2, xi = 0.2, FP32, 1 seed; then the adjoint of the sum of the
                           arrival times at 32 targets (costSensitivity)"""
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
    c, ref = cpu(full)
    r = {"gpu_ms": round(ms, 3), "adjoint_gpu_ms": round(ams, 3), "cpu": c}
    if ref is not None:
        r["speedup_solve"] = round(c["ms"] / ms, 1)
        r["cpu_wall_ms_with_backward"] = c["wall_ms"]
    res["elastica2_192x192x64"] = r
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
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fim_riemann3_128_traffic.json")
    if os.path.exists(tpath) and n == 128:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": "k_fim<Riemann3,float>", "launches": sti["fim_launches"],
        "mean_launch_us": round(mean_launch_ms * 1e3, 2), "unit_of_work": "point-update",
        "bytes_per_point_update": BYTES_PER_POINT_UPDATE, "point_updates": sti["sweeps"] * 64,
        "tile_visits": sti["tile_visits"], "achieved_per_visit": round(achieved_visit, 1),
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

"""Continuous-batching serving of Riemann2 / Riemann3 solves (C ABI hfm_server_*), a Poisson
load generator, and the one-at-a-time baseline it is compared with.

Requests of mixed 2-D / 3-D sizes are admitted by HBM budget, set up (FP64 Selling stencils and
tile graph) on a setup stream while earlier requests iterate, and iterated together: every round
of the band-ordered block fast-iterative method processes the active tiles of all resident
solves in one scheduling and one update launch.
"""
from __future__ import annotations

import ctypes
import math
import time

import numpy as np

from . import _lib
from . import PyExp, Plan

__all__ = ["Server", "make_requests", "run_load", "run_closed_loop", "run_sequential", "latency_summary"]


class Server:
    def __init__(self, hbm_budget_bytes: float = 64e9, max_resident: int = 256, persistent: bool = False):
        """persistent: run each batch of rounds in one persistent launch with grid-wide barriers
        instead of a scheduling and an update launch per round (hfm_server_set_persistent)."""
        self._L = _lib.load()
        h = ctypes.c_void_p()
        if self._L.hfm_server_create(ctypes.byref(h), float(hbm_budget_bytes), int(max_resident)):
            raise PyExp(_lib.last_error())
        self._h = h
        self._keep = {}
        if persistent and self._L.hfm_server_set_persistent(self._h, 1):
            err = _lib.last_error()
            self._L.hfm_server_destroy(self._h)
            self._h = None
            raise PyExp(err)

    def submit(self, model: str, dims, grid_scale: float, metric, seeds, seed_values=None, values_out=None,
               dual: bool = False) -> int:
        """Enqueue one solve.  metric: torch CUDA float64 tensor (kept alive until completion)
        or numpy array (copied).  values_out: optional float64 numpy array of prod(dims)."""
        dim = 2 if model == "Riemann2" else 3
        K = 3 if dim == 2 else 6
        req = _lib.HfmRequest()
        req.model = model.encode()
        d = list(int(n) for n in dims) + [1] * (3 - len(dims))
        req.dims = (ctypes.c_int64 * 3)(*d)
        req.gridScale = float(grid_scale)
        keep = []
        if hasattr(metric, "is_cuda") and metric.is_cuda:
            import torch
            m = metric.to(torch.float64).contiguous()
            torch.cuda.current_stream().synchronize()
            req.metric = m.data_ptr()
            req.metric_on_device = 1
            keep.append(m)
            numel = m.numel()
        else:
            m = np.ascontiguousarray(metric, np.float64)
            req.metric = m.ctypes.data
            req.metric_on_device = 0
            keep.append(m)
            numel = m.size
        req.metric_is_field = int(numel != K)
        req.dual = int(dual)
        s = np.ascontiguousarray(np.asarray(seeds, np.float64).reshape(-1, dim))
        req.nseeds = len(s)
        req.seeds = s.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        keep.append(s)
        if seed_values is not None:
            sv = np.ascontiguousarray(seed_values, np.float64)
            req.seed_values = sv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            keep.append(sv)
        if values_out is not None:
            assert values_out.dtype == np.float64 and values_out.flags["C_CONTIGUOUS"]
            req.values_out = values_out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            keep.append(values_out)
        t = ctypes.c_int64()
        if self._L.hfm_server_submit(self._h, ctypes.byref(req), ctypes.byref(t)):
            raise PyExp(_lib.last_error())
        self._keep[t.value] = keep
        return t.value

    def poll(self, ticket: int) -> bool:
        d = ctypes.c_int()
        if self._L.hfm_server_poll(self._h, ticket, ctypes.byref(d)):
            raise PyExp(_lib.last_error())
        return bool(d.value)

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> dict:
        r = _lib.HfmRequestResult()
        if self._L.hfm_server_wait(self._h, ticket, float(timeout_ms), ctypes.byref(r)):
            raise PyExp(_lib.last_error())
        self._keep.pop(ticket, None)
        return {"status": r.status, "latency_ms": r.latency_ms, "queue_ms": r.queue_ms, "setup_ms": r.setup_ms,
                "rounds": r.rounds, "tile_visits": r.tile_visits, "error": r.error.decode()}

    def stats(self) -> dict:
        st = _lib.HfmServerStats()
        self._L.hfm_server_get_stats(self._h, ctypes.byref(st))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if getattr(self, "_h", None):
            self._L.hfm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


# ---------------------------------------------------------------------------------------------
# Synthetic request stream
# ---------------------------------------------------------------------------------------------
def _smooth_gpu(shape, sigma, gen, torch):
    """Standardised smoothed noise on the GPU: three passes of a periodic box filter of width
    ~sigma*sqrt(12/3) per axis (an approximate Gaussian of std sigma), built from cumulative sums
    so no size-specific kernel has to be compiled."""
    a = torch.randn(shape, generator=gen, device="cuda", dtype=torch.float64)
    r = max(1, int(round(0.5 * (math.sqrt(4.0 * sigma * sigma + 1.0) - 1.0))))  # 3 boxes of 2r+1
    for ax in range(a.dim()):
        n = a.shape[ax]
        rr = min(r, (n - 1) // 2)
        if rr < 1:
            continue
        for _ in range(3):
            pad = torch.cat([a.narrow(ax, n - rr - 1, rr + 1), a, a.narrow(ax, 0, rr)], dim=ax)
            c = torch.cumsum(pad, dim=ax)
            a = (c.narrow(ax, 2 * rr + 1, n) - c.narrow(ax, 0, n)) / (2 * rr + 1)
    a = a - a.mean()
    return a / a.std()


def _rescale(n, lo, hi):
    return lo + (hi - lo) * (n - n.min()) / (n.max() - n.min())


def make_requests(count: int, seed: int = 0, frac2d: float = 0.7, side2d=(64, 1024), side3d=(32, 160),
                  kappa=(1.0, 50.0), seeds=(1, 16), sigma: float = 6.0):
    """The serving mix: 70 % Riemann2 (side log-uniform in [64, 1024]), 30 % Riemann3 (side
    uniform in [32, 160]), 1-16 seeds uniform in [0.05, 0.95]^d, anisotropy kappa uniform in
    [1, 50] (sqrt of the largest eigenvalue ratio).  Metric fields are generated on the GPU."""
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    reqs = []
    for _ in range(count):
        two = rng.uniform() < frac2d
        kap = float(rng.uniform(*kappa))
        ns = int(rng.integers(seeds[0], seeds[1] + 1))
        if two:
            n = int(math.floor(math.exp(rng.uniform(math.log(side2d[0]), math.log(side2d[1])))))
            th = math.pi * _smooth_gpu((n, n), sigma, gen, torch)
            r = _rescale(_smooth_gpu((n, n), sigma, gen, torch), 1.0, kap)
            c, s = torch.cos(th), torch.sin(th)
            m = torch.stack([c * c + s * s * r * r, c * s * (1 - r * r), s * s + c * c * r * r], dim=-1)
            dims = (n, n)
            model = "Riemann2"
        else:
            n = int(rng.integers(side3d[0], side3d[1] + 1))
            shp = (n, n, n)
            al, be, ga = (math.pi * _smooth_gpu(shp, sigma, gen, torch) for _ in range(3))
            a = _rescale(_smooth_gpu(shp, sigma, gen, torch), 1.0, kap)
            b = _rescale(_smooth_gpu(shp, sigma, gen, torch), 1.0, kap)
            m = _metric3_torch(al, be, ga, a, b, torch)
            dims = shp
            model = "Riemann3"
        sd = rng.uniform(0.05, 0.95, size=(ns, len(dims)))
        reqs.append({"model": model, "dims": dims, "gridScale": 1.0 / n, "metric": m.contiguous(), "seeds": sd,
                     "kappa": kap})
    torch.cuda.synchronize()
    return reqs


def _metric3_torch(al, be, ga, a, b, torch):
    def rz(t):
        c, s = torch.cos(t), torch.sin(t)
        z, o = torch.zeros_like(t), torch.ones_like(t)
        return torch.stack([torch.stack([c, -s, z], -1), torch.stack([s, c, z], -1), torch.stack([z, z, o], -1)], -2)

    def ry(t):
        c, s = torch.cos(t), torch.sin(t)
        z, o = torch.zeros_like(t), torch.ones_like(t)
        return torch.stack([torch.stack([c, z, s], -1), torch.stack([z, o, z], -1), torch.stack([-s, z, c], -1)], -2)

    R = rz(al) @ ry(be) @ rz(ga)
    lam = torch.stack([torch.ones_like(a), a * a, b * b], -1)
    M = (R * lam.unsqueeze(-2)) @ R.transpose(-1, -2).contiguous()
    return torch.stack([M[..., 0, 0], M[..., 1, 0], M[..., 1, 1], M[..., 2, 0], M[..., 2, 1], M[..., 2, 2]], -1)


def latency_summary(lat_ms):
    a = np.asarray(lat_ms, float)
    return {"p50_ms": float(np.percentile(a, 50)), "p99_ms": float(np.percentile(a, 99)),
            "mean_ms": float(a.mean()), "max_ms": float(a.max())}


def run_load(server: Server, requests, qps: float, seed: int = 1, collect_values: bool = False):
    """Submit requests at Poisson arrival times of rate qps (qps <= 0: all at once); wait for all.
    Returns solves/s over the makespan and the latency percentiles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gaps = rng.exponential(1.0 / qps, size=len(requests)) if qps > 0 else np.zeros(len(requests))
    outs = [np.empty(int(np.prod(r["dims"])), np.float64) if collect_values else None for r in requests]
    tickets = []
    t0 = time.perf_counter()
    next_t = t0
    for r, g, o in zip(requests, gaps, outs):
        next_t += g
        while True:
            now = time.perf_counter()
            if now >= next_t:
                break
            time.sleep(min(next_t - now, 0.002))
        tickets.append(server.submit(r["model"], r["dims"], r["gridScale"], r["metric"], r["seeds"], values_out=o))
    res = [server.wait(t) for t in tickets]
    t1 = time.perf_counter()
    ok = [x for x in res if x["status"] == 0]
    out = {"qps_offered": qps, "requests": len(requests), "completed": len(ok),
           "solves_per_s": len(ok) / (t1 - t0), "makespan_s": t1 - t0}

    out.update(latency_summary([x["latency_ms"] for x in ok]))
    out["queue_p50_ms"] = float(np.percentile([x["queue_ms"] for x in ok], 50))
    if collect_values:
        out["values"] = [o.reshape(r["dims"]) for o, r in zip(outs, requests)]
    return out


def run_closed_loop(server: Server, requests, concurrency: int = 128, rounds: int = 3):
    """Closed-loop load: keep `concurrency` requests in flight, submitting the next one (cycling
    through `requests`, `rounds` times) whenever one completes.  The sustained rate counts the
    completions between the concurrency-th and the (total - concurrency)-th, a window in which
    the server always holds `concurrency` requests and whose mix is that of the stream."""
    total = len(requests) * rounds
    inflight = {}
    done_times = []
    lat = []
    nxt = 0
    t0 = time.perf_counter()

    def submit_next():
        nonlocal nxt
        r = requests[nxt % len(requests)]
        inflight[server.submit(r["model"], r["dims"], r["gridScale"], r["metric"], r["seeds"])] = nxt
        nxt += 1

    while nxt < min(concurrency, total):
        submit_next()
    while inflight:
        finished = [t for t in list(inflight) if server.poll(t)]
        if not finished:
            time.sleep(0.0002)
            continue
        now = time.perf_counter()
        for t in finished:
            res = server.wait(t)
            del inflight[t]
            if res["status"] != 0:
                raise PyExp("request failed: " + res["error"])
            done_times.append(now)
            lat.append(res["latency_ms"])
            if nxt < total:
                submit_next()
    t1 = time.perf_counter()
    done_times = np.sort(np.asarray(done_times))
    a, b = concurrency, total - concurrency
    out = {"requests": total, "concurrency": concurrency, "makespan_s": t1 - t0,
           "solves_per_s": len(done_times) / (t1 - t0)}
    if b > a + 8:
        out["sustained_solves_per_s"] = (b - a) / max(done_times[b - 1] - done_times[a - 1], 1e-9)
    out.update(latency_summary(lat))
    return out


def run_sequential(requests, collect_values: bool = False):
    """One-at-a-time plain GPU solves of the same requests (a plan per request, FP32 update,
    values copied to the host), as a client would issue them without the server."""
    import torch
    vals = []
    lat = []
    t0 = time.perf_counter()
    for r in requests:
        ts = time.perf_counter()
        p = Plan(r["model"], r["dims"], r["gridScale"], precision=32)
        p.set_metric(r["metric"])
        p.solve(r["seeds"])
        v = p.values().cpu().numpy()
        lat.append((time.perf_counter() - ts) * 1e3)
        if collect_values:
            vals.append(v)
        del p
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    out = {"requests": len(requests), "solves_per_s": len(requests) / (t1 - t0), "makespan_s": t1 - t0}
    out.update(latency_summary(lat))
    if collect_values:
        out["values"] = vals
    return out


def main(argv=None):
    """python -m hamiltonfastmarching_amd.serving [--requests N] [--concurrency C] [--rounds R]
    [--burst] [--seed S] [--persistent]: the closed-loop sustained rate (or a burst) on the request
    mix, as JSON."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--concurrency", type=int, default=128)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--burst", action="store_true")
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--persistent", action="store_true", help="persistent round kernel (grid-wide barriers)")
    a = ap.parse_args(argv)
    reqs = make_requests(a.requests, seed=a.seed)
    srv = Server(hbm_budget_bytes=64e9, max_resident=256, persistent=a.persistent)
    try:
        run_load(srv, reqs[:16], qps=0)
        out = run_load(srv, reqs, qps=0) if a.burst else run_closed_loop(srv, reqs, a.concurrency, a.rounds)
    finally:
        srv.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

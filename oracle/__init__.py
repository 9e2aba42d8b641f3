"""Checker for the GPU library: a CPU restatement of HFM's hot path, plus a runner for the
reference's own FileHFM executables where they have been built from source.

TEST INFRASTRUCTURE ONLY.  Nothing in ``hamiltonfastmarching_amd`` imports this package.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it, and
only as the checker or the timed CPU baseline -- never as the computation being measured.

* ``libhfm_oracle.so`` (``oracle/hfm_oracle.cpp``): FP64 restatement of the Selling reduction,
  the curvature stencils, the heap Fast-Marching loop, the geodesic flow and the backward
  differentiation.  Built by ``make -C oracle``.
* ``oracle/_ref/FileHFM_<Model>``: the reference front-end ``Interfaces/FileHFM/FileHFM.cpp``
  compiled from its sources by ``make -C oracle ref`` (only where ``/root/reference`` exists).
  It is driven through the reference's file format (``JMM_CPPLibs/Output/FileIO.hxx:7-132``:
  ``<prefix>_Format.txt`` + ``<prefix>_Data.dat`` of float64), written by ``write_rules`` here.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile
from functools import reduce
from operator import mul

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libhfm_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")

MODELS = {"Riemann2": 0, "Riemann3": 1, "ReedsSheppForward2": 2, "Elastica2": 3}
MODEL_DIM = {"Riemann2": 2, "Riemann3": 3, "ReedsSheppForward2": 3, "Elastica2": 3}

_lib = None


def build(ref_sources: str | None = "/root/reference") -> None:
    """Compile the restatement, and the reference binaries when their sources are present."""
    subprocess.check_call(["make", "-s", "-C", HERE])
    if ref_sources and os.path.isdir(ref_sources):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref", f"REF={ref_sources}"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d, i32, i64 = ctypes.c_double, ctypes.c_int, ctypes.c_int64
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_riemann_stencils.argtypes = [i32, P(i32), P(d), P(d), i32, i32, P(ctypes.c_int8), P(d)]
        L.oracle_curvature_stencils.argtypes = [i32, i32, d, d, d, P(ctypes.c_int8), P(d), P(i32)]
        L.oracle_solve.argtypes = [i32, P(i32), P(d), P(d), P(d), i32, i32, d, d, P(d), i32, i32, P(d), P(d),
                                   P(d), P(ctypes.c_uint64), P(d), P(d)]
        L.oracle_solve_backward.argtypes = [i32, P(i32), P(d), P(d), P(d), i32, i32, d, d, P(d), i32, i32,
                                            P(d), P(d), i32, P(d), P(d), P(d), P(d), P(i64), P(d), i32, P(i32)]
        L.oracle_geodesics.argtypes = [i32, P(i32), P(d), P(d), P(d), i32, i32, d, d, P(d), i32, i32, P(d), P(d),
                                       i32, P(d), d, d, d, P(d), ctypes.c_long, P(i64), P(ctypes.c_long)]
        L.oracle_forward_variation.argtypes = [i32, P(i32), P(d), P(d), P(d), i32, i32, d, d, P(d), i32, i32, P(d),
                                               P(d), i32, P(d), P(d), P(d), P(d)]
        L.oracle_set_walls.argtypes = [ctypes.c_long, P(ctypes.c_ubyte), i32]
        L.oracle_set_walls.restype = None
        L.oracle_set_order.argtypes = [i32]
        L.oracle_set_order.restype = None
        L.oracle_last_dips.restype = ctypes.c_long
        L.oracle_set_stopping.argtypes = [ctypes.c_long, d, i32, P(ctypes.c_long), i32, P(ctypes.c_long), i32,
                                          P(ctypes.c_long), i32, P(ctypes.c_long)]
        L.oracle_set_stopping.restype = None
        L.oracle_last_naccepted.restype = ctypes.c_long
        L.oracle_set_factoring.argtypes = [i32, i32, P(i32), P(i32), P(d), P(d), P(ctypes.c_ubyte)]
        L.oracle_set_factoring.restype = None
        L.oracle_set_geodesic_solver.argtypes = [i32, d]
        L.oracle_set_geodesic_solver.restype = None
        L.oracle_last_geodesic_failures.restype = ctypes.c_long
        L.oracle_geodesics_given.argtypes = [i32, P(i32), P(d), P(d), P(d), i32, i32, d, d, P(d), i32, P(d),
                                             P(ctypes.c_uint64), i32, P(d), d, d, d, P(d), ctypes.c_long, P(i64),
                                             P(ctypes.c_long), i32, P(d)]
        _lib = L
    return _lib


def _ptr(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _err():
    return lib().oracle_last_error().decode()


def riemann_stencils(dims, h, metric, dual=False):
    """Selling stencils: (offsets int8 [N,K,d], weights float64 [N,K])."""
    dims = [int(n) for n in dims]
    dim = len(dims)
    K = 3 if dim == 2 else 6
    h = np.broadcast_to(np.asarray(h, np.float64), (dim,)).copy()
    metric = np.ascontiguousarray(metric, np.float64)
    N = reduce(mul, dims)
    is_field = metric.size != K
    off = np.zeros((N, K, dim), np.int8)
    w = np.zeros((N, K))
    dimsa = np.array(dims, np.int32)
    r = lib().oracle_riemann_stencils(dim, _ptr(dimsa, ctypes.c_int), _ptr(h, ctypes.c_double),
                                      _ptr(metric, ctypes.c_double), int(is_field), int(dual),
                                      _ptr(off, ctypes.c_int8), _ptr(w, ctypes.c_double))
    if r:
        raise RuntimeError(_err())
    return off, w


def curvature_stencils(model, ntheta, h, xi, eps=0.1):
    """Shared per-theta stencils: (offsets [ntheta,n,3], weights [ntheta,n], symmetric [n])."""
    n = 4 if model == "ReedsSheppForward2" else 27
    off = np.zeros((ntheta, n, 3), np.int8)
    w = np.zeros((ntheta, n))
    sym = np.zeros(n, np.int32)
    r = lib().oracle_curvature_stencils(MODELS[model], int(ntheta), float(h), float(xi), float(eps),
                                        _ptr(off, ctypes.c_int8), _ptr(w, ctypes.c_double),
                                        _ptr(sym, ctypes.c_int))
    if r:
        raise RuntimeError(_err())
    return off, w, sym.astype(bool)


def _common_args(params):
    model = params["model"]
    dims = [int(n) for n in np.asarray(params["dims"]).ravel()]
    dim = MODEL_DIM[model]
    if len(dims) != dim:
        raise ValueError("dims has the wrong length")
    N = reduce(mul, dims)
    if "gridScales" in params:
        h = np.asarray(params["gridScales"], np.float64).ravel().copy()
    else:
        h = np.full(dim, float(params["gridScale"]))
    origin = params.get("origin")
    origin = None if origin is None else np.asarray(origin, np.float64).ravel().copy()
    metric = None
    dual = 0
    metric_is_field = 0
    cost = None
    cost_is_field = 0
    if model in ("Riemann2", "Riemann3"):
        if "dualMetric" in params:
            metric, dual = params["dualMetric"], 1
        else:
            metric = params["metric"]
        metric = np.ascontiguousarray(metric, np.float64)
        metric_is_field = int(metric.ndim > 1)
    else:
        if "speed" in params:
            cost = 1.0 / np.ascontiguousarray(params["speed"], np.float64)
        else:
            cost = np.ascontiguousarray(params["cost"], np.float64)
        cost = np.ascontiguousarray(np.broadcast_to(cost.reshape(cost.shape + (1,) * (3 - cost.ndim))
                                                    if 0 < cost.ndim < 3 else cost, tuple(dims)))
        cost_is_field = 1
    seeds = np.ascontiguousarray(np.asarray(params["seeds"], np.float64).reshape(-1, dim))
    sv = params.get("seedValues")
    sv = None if sv is None else np.ascontiguousarray(sv, np.float64).ravel()
    xi = float(params.get("xi", 1.0))
    eps = float(params.get("eps", 0.1))
    _set_walls(params, dims, model)
    lib().oracle_set_order(int(float(params.get("order", 1))))
    _set_stopping(params, dims, model, h, origin)
    seeds, sv = _set_factoring(params, model, dims, h, origin, dual, seeds, sv)
    return model, dims, dim, N, h, origin, metric, metric_is_field, dual, cost, cost_is_field, seeds, sv, xi, eps


def _stop_indices(params, key, dims, model, h, origin, unoriented):
    """IndexFromPoint(ADim(p)) of the targets, periodized, out-of-range ones dropped
    (ExtraAlgorithms/CommonStoppingCriteria.h:115-151); row-major indices, (x, y) ones when unoriented."""
    if key not in params:
        return np.zeros(0, np.int64)
    pd = 2 if unoriented else len(dims)
    pts = np.asarray(params[key], np.float64).reshape(-1, pd)
    bundle = model in ("ReedsSheppForward2", "Elastica2")
    out = []
    for p in pts:
        idx = []
        ok = True
        for i in range(pd):
            if bundle and i == 2:
                ds = 2 * np.pi / dims[2]
                v = int(np.floor((p[i] + ds / 2) / ds)) % dims[2]
            else:
                v = int(np.floor((p[i] - (0. if origin is None else origin[i])) / h[i]))
                ok = ok and 0 <= v < dims[i]
            idx.append(v)
        if ok:
            out.append(np.ravel_multi_index(tuple(idx), tuple(dims[:pd])))
    return np.asarray(out, np.int64)


def _set_stopping(params, dims, model, h, origin):
    bundle = model in ("ReedsSheppForward2", "Elastica2")
    lists = [_stop_indices(params, k, dims, model, h, origin, False)
             for k in ("stopWhenAnyAccepted", "stopWhenAllAccepted")]
    lists += [_stop_indices(params, k, dims, model, h, origin, True) if bundle else np.zeros(0, np.int64)
              for k in ("stopWhenAnyAccepted_Unoriented", "stopWhenAllAccepted_Unoriented")]
    lists = [np.ascontiguousarray(a, dtype=np.int64) for a in lists]
    nmax = int(float(params["nMaxAccepted"])) if "nMaxAccepted" in params else -1
    dist = float(params["stopAtDistance"]) if "stopAtDistance" in params else float("inf")
    args = []
    for a in lists:
        args += [len(a), a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))]
    lib().oracle_set_stopping(nmax, dist, *args)


_last_exports: dict = {}


def _set_factoring(params, model, dims, h, origin, dual, seeds, sv):
    """Seed spreading and the static factor (oracle/factoring.py) for the next solve: returns the
    seeds and values the solver starts from."""
    from . import factoring as fac
    global _last_exports
    _last_exports = {}
    keys = ("factoringRadius", "factoringValues", "seedRadius")
    f = None
    if any(k in params for k in keys):
        raw = params.get("dualMetric", params.get("metric")) if model in ("Riemann2", "Riemann3") else None
        seeds, sv, f, _last_exports = fac.setup(params, model, dims, h, origin, raw, dual, seeds, sv)
        seeds = np.ascontiguousarray(seeds, np.float64)
        sv = np.ascontiguousarray(sv, np.float64)
    if f is None:
        lib().oracle_set_factoring(len(dims), 0, None, None, None, None, None)
    else:
        sub, bd, shift, vals, grads, mask = f
        bda = np.array(bd, np.int32)
        sha = np.array(shift, np.int32)
        vals = np.ascontiguousarray(vals, np.float64)
        grads = np.ascontiguousarray(grads, np.float64)
        mask = None if mask is None else np.ascontiguousarray(mask, np.uint8)
        lib().oracle_set_factoring(len(dims), int(sub), _ptr(bda, ctypes.c_int), _ptr(sha, ctypes.c_int),
                                   _ptr(vals, ctypes.c_double), _ptr(grads, ctypes.c_double),
                                   _ptr(mask, ctypes.c_ubyte))
    return seeds, sv


def last_naccepted() -> int:
    """nAccepted of the last solve (ExtraAlgorithms/CommonStoppingCriteria.h:162-165)."""
    return int(lib().oracle_last_naccepted())


def last_dips() -> int:
    """Pops of the last solve whose key was below the previous pop's key."""
    return int(lib().oracle_last_dips())


def grid_walls(params, dims, model):
    """The `walls` field as one byte per point (nonzero = wall), or None: a constant, a full
    field, or for the curvature models an (x, y) field or a theta field broadcast
    (GetIntegralField, HFMInterface.hxx:33-42)."""
    if "walls" not in params:
        return None
    w = np.asarray(params["walls"], np.float64)
    if w.ndim == 0:
        full = np.full(tuple(dims), w != 0)
    elif w.ndim == len(dims):
        full = w != 0
    elif model not in ("Riemann2", "Riemann3") and w.ndim == 2:
        full = np.broadcast_to((w != 0)[:, :, None], tuple(dims))
    elif model not in ("Riemann2", "Riemann3") and w.ndim == 1:
        full = np.broadcast_to((w != 0)[None, None, :], tuple(dims))
    else:
        raise RuntimeError("Field walls has incorrect depth.")
    if full.shape != tuple(dims):
        raise RuntimeError("Error : Field walls has inconsistent dimensions.")
    return np.ascontiguousarray(full, np.uint8).ravel()


def _set_walls(params, dims, model):
    w = grid_walls(params, dims, model)
    refine = int(float(params.get("refineStencilAtWallBoundary", 0)) != 0)
    refine |= 2 if params.get("_plant_no_wall_walk") else 0  # planted mistake, tests only
    if w is None:
        lib().oracle_set_walls(0, None, 0)
    else:
        lib().oracle_set_walls(w.size, _ptr(w, ctypes.c_ubyte), refine)
        _set_walls.keep = w  # the library copies the bytes; keep them alive across the call anyway


def solve(params, flow=False, active=False):
    """FP64 Fast Marching on an HFM-style input dict.  Returns a dict with 'values'
    (and 'activeNeighs', 'geodesicFlow', 'StencilCPUTime', 'FMCPUTime')."""
    (model, dims, dim, N, h, origin, metric, mif, dual, cost, cif, seeds, sv, xi, eps) = _common_args(params)
    values = np.zeros(N)
    act = np.zeros(N, np.uint64) if active else None
    fl = np.zeros((N, dim)) if flow else None
    tm = np.zeros(2)
    dimsa = np.array(dims, np.int32)
    r = lib().oracle_solve(MODELS[model], _ptr(dimsa, ctypes.c_int), _ptr(h, ctypes.c_double),
                           _ptr(origin, ctypes.c_double), _ptr(metric, ctypes.c_double), mif, dual, xi, eps,
                           _ptr(cost, ctypes.c_double), cif, len(seeds), _ptr(seeds, ctypes.c_double),
                           _ptr(sv, ctypes.c_double), _ptr(values, ctypes.c_double),
                           _ptr(act, ctypes.c_uint64), _ptr(fl, ctypes.c_double), _ptr(tm, ctypes.c_double))
    if r:
        raise RuntimeError(_err())
    out = {"values": values.reshape(dims), "StencilCPUTime": tm[0], "FMCPUTime": tm[1],
           "nAccepted": last_naccepted(), **_last_exports}
    if active:
        out["activeNeighs"] = act.reshape(dims)
    if flow:
        out["geodesicFlow"] = fl.reshape(tuple(dims) + (dim,))
    return out


def solve_backward(params, targets, weights=None):
    """FM + BackwardVariation for one group of targets: (values, costSensitivity, seedSens)."""
    (model, dims, dim, N, h, origin, metric, mif, dual, cost, cif, seeds, sv, xi, eps) = _common_args(params)
    targets = np.ascontiguousarray(np.asarray(targets, np.float64).reshape(-1, dim))
    tw = np.ones(len(targets)) if weights is None else np.ascontiguousarray(weights, np.float64)
    values = np.zeros(N)
    sens = np.zeros(N)
    nmax = 1024
    si = np.zeros(nmax, np.int64)
    sw = np.zeros(nmax)
    ns = ctypes.c_int(0)
    dimsa = np.array(dims, np.int32)
    r = lib().oracle_solve_backward(MODELS[model], _ptr(dimsa, ctypes.c_int), _ptr(h, ctypes.c_double),
                                    _ptr(origin, ctypes.c_double), _ptr(metric, ctypes.c_double), mif, dual,
                                    xi, eps, _ptr(cost, ctypes.c_double), cif, len(seeds),
                                    _ptr(seeds, ctypes.c_double), _ptr(sv, ctypes.c_double), len(targets),
                                    _ptr(targets, ctypes.c_double), _ptr(tw, ctypes.c_double),
                                    _ptr(values, ctypes.c_double), _ptr(sens, ctypes.c_double),
                                    _ptr(si, ctypes.c_int64), _ptr(sw, ctypes.c_double), nmax, ctypes.byref(ns))
    if r:
        raise RuntimeError(_err())
    n = min(ns.value, nmax)
    return values.reshape(dims), sens.reshape(dims), (si[:n].copy(), sw[:n].copy())


def geodesics(params, tips, step=0.25, threshold=0.001, volume_bound=-1.0):
    """FM + the geodesic solver from each tip: (points (M x dim), lengths (ntips)).  The discrete
    solver, or the ODE solver when params['geodesicSolver'] is "ODE" (its step and causality
    tolerance from geodesicStep / geodesicCausalityTolerance, defaults 0.25 and 4)."""
    (model, dims, dim, N, h, origin, metric, mif, dual, cost, cif, seeds, sv, xi, eps) = _common_args(params)
    ode = params.get("geodesicSolver", "Discrete") == "ODE"
    lib().oracle_set_geodesic_solver(int(ode), float(params.get("geodesicCausalityTolerance", 4.0)))
    if ode:
        step = float(params.get("geodesicStep", step))
    tips = np.ascontiguousarray(np.asarray(tips, np.float64).reshape(-1, dim))
    lengths = np.zeros(len(tips), np.int64)
    npts = ctypes.c_long(0)
    dimsa = np.array(dims, np.int32)
    cap = 1 << 16
    while True:
        pts = np.zeros((cap, dim))
        r = lib().oracle_geodesics(MODELS[model], _ptr(dimsa, ctypes.c_int), _ptr(h, ctypes.c_double),
                                   _ptr(origin, ctypes.c_double), _ptr(metric, ctypes.c_double), mif, dual,
                                   xi, eps, _ptr(cost, ctypes.c_double), cif, len(seeds),
                                   _ptr(seeds, ctypes.c_double), _ptr(sv, ctypes.c_double), len(tips),
                                   _ptr(tips, ctypes.c_double), float(step), float(threshold),
                                   float(volume_bound), _ptr(pts, ctypes.c_double), cap,
                                   _ptr(lengths, ctypes.c_int64), ctypes.byref(npts))
        if r:
            raise RuntimeError(_err())
        if npts.value <= cap:
            lib().oracle_set_geodesic_solver(0, 4.0)
            return pts[:npts.value].copy(), lengths.copy()
        cap = npts.value


def last_geodesic_failures() -> int:
    """Tips of the last ODE run that stalled (the reference reports them as failed)."""
    return int(lib().oracle_last_geodesic_failures())


def forward_variation(params, cost_variation=None, seed_value_variation=None):
    """FM + ForwardVariation: (values, valueVariation of shape dims + (nVar,)).
    cost_variation: dims + (nVar,); seed_value_variation: (nVar, nseeds)."""
    (model, dims, dim, N, h, origin, metric, mif, dual, cost, cif, seeds, sv, xi, eps) = _common_args(params)
    cvar = svar = None
    nvar = 0
    if cost_variation is not None:
        cvar = np.ascontiguousarray(np.asarray(cost_variation, np.float64).reshape(N, -1))
        nvar = cvar.shape[1]
    if seed_value_variation is not None:
        given = np.asarray(params["seeds"], np.float64).reshape(-1, dim)
        svar = np.asarray(seed_value_variation, np.float64).reshape(-1, len(given))
        if len(given) != len(seeds) or not np.array_equal(given, seeds):
            # the solver's seeds were spread: the variations sit at the given seeds' grid points
            # (the last one written wins), every other solver seed varies by 0 (FirstVariation.h:106-119)
            def cell(pt):  # IndexFromPoint(ADim(p)) on a Riemann grid (spreading is Riemann only)
                return tuple(int(np.floor((pt[i] - (0. if origin is None else origin[i])) / h[i])) for i in range(dim))
            at = {}
            for i, pt in enumerate(given):
                at[cell(pt)] = i
            sv2 = np.zeros((svar.shape[0], len(seeds)))
            for j, pt in enumerate(seeds):
                i = at.get(cell(pt))
                if i is not None:
                    sv2[:, j] = svar[:, i]
            svar = sv2
        svar = np.ascontiguousarray(svar)
        nvar = max(nvar, svar.shape[0])
    values = np.zeros(N)
    out = np.zeros((N, nvar))
    dimsa = np.array(dims, np.int32)
    r = lib().oracle_forward_variation(MODELS[model], _ptr(dimsa, ctypes.c_int), _ptr(h, ctypes.c_double),
                                       _ptr(origin, ctypes.c_double), _ptr(metric, ctypes.c_double), mif, dual,
                                       xi, eps, _ptr(cost, ctypes.c_double), cif, len(seeds),
                                       _ptr(seeds, ctypes.c_double), _ptr(sv, ctypes.c_double), nvar,
                                       _ptr(cvar, ctypes.c_double), _ptr(svar, ctypes.c_double),
                                       _ptr(values, ctypes.c_double), _ptr(out, ctypes.c_double))
    if r:
        raise RuntimeError(_err())
    return values.reshape(dims), out.reshape(tuple(dims) + (nvar,))


def geodesics_given(params, values, active, tips, step=0.25, threshold=0.001, volume_bound=-1.0):
    """The geodesic solver on given values / activeNeighs (e.g. another solver's): the discrete
    one, or the ODE one when params['geodesicSolver'] is "ODE" (as in geodesics; the seeds'
    distance from params['seeds'])."""
    (model, dims, dim, N, h, origin, metric, mif, dual, cost, cif, seeds, sv, xi, eps) = _common_args(params)
    ode = params.get("geodesicSolver", "Discrete") == "ODE"
    lib().oracle_set_geodesic_solver(int(ode), float(params.get("geodesicCausalityTolerance", 4.0)))
    if ode:
        step = float(params.get("geodesicStep", step))
    tips = np.ascontiguousarray(np.asarray(tips, np.float64).reshape(-1, dim))
    vals = np.ascontiguousarray(np.asarray(values, np.float64).ravel())
    act = np.ascontiguousarray(np.asarray(active).astype(np.uint64).ravel())
    assert vals.size == N and act.size == N
    lengths = np.zeros(len(tips), np.int64)
    npts = ctypes.c_long(0)
    dimsa = np.array(dims, np.int32)
    cap = 1 << 16
    while True:
        pts = np.zeros((cap, dim))
        r = lib().oracle_geodesics_given(MODELS[model], _ptr(dimsa, ctypes.c_int), _ptr(h, ctypes.c_double),
                                         _ptr(origin, ctypes.c_double), _ptr(metric, ctypes.c_double), mif, dual,
                                         xi, eps, _ptr(cost, ctypes.c_double), cif, _ptr(vals, ctypes.c_double),
                                         _ptr(act, ctypes.c_uint64), len(tips), _ptr(tips, ctypes.c_double),
                                         float(step), float(threshold), float(volume_bound),
                                         _ptr(pts, ctypes.c_double), cap, _ptr(lengths, ctypes.c_int64),
                                         ctypes.byref(npts), len(seeds), _ptr(seeds, ctypes.c_double))
        if r:
            lib().oracle_set_geodesic_solver(0, 4.0)
            raise RuntimeError(_err())
        if npts.value <= cap:
            lib().oracle_set_geodesic_solver(0, 4.0)
            return pts[:npts.value].copy(), lengths.copy()
        cap = npts.value


# ------------------------------------------------------------------------------------------
# Reference executables, driven through the reference's file format.
# ------------------------------------------------------------------------------------------
def ref_binary(model: str) -> str:
    return os.path.join(REF_DIR, f"FileHFM_{model}")


def ref_available(model: str) -> bool:
    return os.path.exists(ref_binary(model))


def write_rules(params: dict, prefix: str) -> None:
    """Write an input dict in the FileHFM format (FileIO.hxx:7-95): per key, the key name, the
    number of dims (-1 for strings, 0 for scalars), the dims, a blank line; data as float64."""
    data = []
    with open(prefix + "_Format.txt", "w") as f:
        for key, val in params.items():
            if isinstance(val, str):
                f.write(f"{key}\n-1\n{val}\n\n")
            elif np.isscalar(val):
                f.write(f"{key}\n0\n\n")
                data.append(np.array([float(val)]))
            else:
                a = np.asarray(val, np.float64)
                f.write(f"{key}\n{a.ndim}\n")
                for n in a.shape:
                    f.write(f"{n}\n")
                f.write("\n")
                data.append(a.ravel())
    np.concatenate(data).astype("<f8").tofile(prefix + "_Data.dat")


def read_rules(prefix: str) -> dict:
    """Read the FileHFM output pair (FileIO.hxx:97-132)."""
    data = np.fromfile(prefix + "_Data.dat", dtype="<f8")
    out, pos = {}, 0
    with open(prefix + "_Format.txt") as f:
        while True:
            key = f.readline().strip()
            if not key:
                break
            nd = int(f.readline())
            if nd == -1:
                out[key] = f.readline().rstrip("\n")
            elif nd == 0:
                out[key] = float(data[pos])
                pos += 1
            else:
                dims = [int(f.readline()) for _ in range(nd)]
                size = reduce(mul, dims, 1)
                out[key] = data[pos:pos + size].reshape(dims).copy()
                pos += size
            f.readline()
    return out


def run_reference(params: dict, timeout: float = 600.0) -> dict:
    """Run the reference FileHFM executable of params['model'] on params; returns its outputs."""
    model = params["model"]
    exe = ref_binary(model)
    if not os.path.exists(exe):
        raise RuntimeError(f"reference binary not built: {exe}")
    with tempfile.TemporaryDirectory() as td:
        ip, op = os.path.join(td, "input"), os.path.join(td, "output")
        write_rules({k: v for k, v in params.items()}, ip)
        proc = subprocess.run([exe, ip, op], capture_output=True, text=True, timeout=timeout)
        out = read_rules(op) if os.path.exists(op + "_Format.txt") else {}
        out["retcode"] = proc.returncode
        out["log"] = proc.stdout + proc.stderr
    return out


def parse_stencils(s: str, dim: int):
    """Parse the 'stencils' string printed for the getStencils key (HFMInterface.hxx:360-375,
    EulerianStencil.hxx:76-87): returns a list of (index, [(weight, offset), ...]) in print order."""
    import re
    out = []
    # each entry: {{i,j[,k]},{forward...,symmetric...}}
    num = r"-?\d+"
    flt = r"[-+0-9.eE]+|inf|nan"
    pat_entry = re.compile(r"\{\{(" + num + r"(?:," + num + r")*)\},(\{.*?\}\})\}")
    pat_diff = re.compile(r"\{(" + flt + r"),\{(" + num + r"(?:," + num + r")*)\}\}")
    for m in pat_entry.finditer(s):
        idx = tuple(int(t) for t in m.group(1).split(","))
        diffs = [(float(d.group(1)), tuple(int(t) for t in d.group(2).split(",")))
                 for d in pat_diff.finditer(m.group(2))]
        out.append((idx, diffs))
    return out

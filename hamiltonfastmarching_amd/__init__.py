"""MI355X (gfx950) Fast Marching: Python surface of libhfm_amd.so.

Mirrors the reference's Python interface for its hot path:

* ``HFMIO`` -- the pybind class ``HFM_<Model>.HFMIO`` (Interfaces/PythonHFM/src/main.cpp:11-47):
  ``run``, ``get/set_string``, ``get/set_scalar``, ``get/set_array``, ``computed_keys``,
  ``has_field``, ``erase_field``.  Errors raise ``PyExp`` like the registered
  ``JMMCppException`` (main.cpp:31).  The default arrayOrdering is ``YXZ_RowMajor`` as in
  PythonIO (JMM_CPPLibs/JMM_CPPLibs/Output/PythonIO.h:20); one library serves every model,
  chosen by the ``model`` key.
* ``run(params)`` -- dictionary in, dictionary of computed keys out, the in-process analogue of
  ``FileIO.WriteCallRead`` (JMM_CPPLibs/FileIO/FileIO.py:54-73) with FileHFM's RowMajor default.
* ``Plan`` -- device-resident solver over torch CUDA tensors (hfm_plan_* of the C ABI).
* ``EikonalSolve`` -- ``torch.autograd.Function``: arrival times at target points as a function of
  the cost field, backward = the adjoint (FirstVariation::BackwardVariation), forward-mode
  ``jvp`` = the forward variation (FirstVariation::ForwardVariation).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

__all__ = ["HFMIO", "PyExp", "run", "Plan", "EikonalSolve", "MODELS"]

MODELS = ("Riemann2", "Riemann3", "ReedsSheppForward2", "Elastica2")
MODEL_DIM = {"Riemann2": 2, "Riemann3": 3, "ReedsSheppForward2": 3, "Elastica2": 3, "Elastica2_5": 3}


class PyExp(RuntimeError):
    """Raised on any failure of the library (mirrors the pybind exception of the reference)."""


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class HFMIO:
    """Key/value IO object; ``run()`` executes the solve on the GPU."""

    def __init__(self, params: dict | None = None, default_ordering: str = "YXZ_RowMajor"):
        self._L = _lib.load()
        self._io = self._L.hfm_io_create()
        if not self._io:
            raise PyExp("hfm_io_create failed")
        if default_ordering:
            self.set_string("arrayOrdering", default_ordering)
        for k, v in (params or {}).items():
            self[k] = v

    def __del__(self):
        io = getattr(self, "_io", None)
        if io:
            self._L.hfm_io_destroy(io)
            self._io = None

    def _err(self):
        return self._L.hfm_io_last_error(self._io).decode()

    # setters -----------------------------------------------------------------------------
    def set_string(self, key: str, value: str) -> None:
        if self._L.hfm_io_set_string(self._io, key.encode(), value.encode()):
            raise PyExp(self._err())

    def set_scalar(self, key: str, value: float) -> None:
        if self._L.hfm_io_set_scalar(self._io, key.encode(), float(value)):
            raise PyExp(self._err())

    def set_array(self, key: str, value) -> None:
        """Copy an array into the IO store (FP64, as the reference stores every numerical field);
        float32 arrays go through the FP32 setter, without an FP64 copy on the Python side."""
        if isinstance(value, np.ndarray) and value.dtype == np.float32:
            a = np.ascontiguousarray(value)
            shape = (ctypes.c_int64 * max(a.ndim, 1))(*a.shape)
            rc = self._L.hfm_io_set_array_f32(self._io, key.encode(), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                              a.ndim, shape)
        else:
            a = np.ascontiguousarray(value, dtype=np.float64)
            shape = (ctypes.c_int64 * max(a.ndim, 1))(*a.shape)
            rc = self._L.hfm_io_set_array(self._io, key.encode(), _dptr(a), a.ndim, shape)
        if rc:
            raise PyExp(self._err())

    def __setitem__(self, key, value):
        if isinstance(value, str):
            self.set_string(key, value)
        elif np.isscalar(value):
            self.set_scalar(key, value)
        else:
            self.set_array(key, value)

    # getters -----------------------------------------------------------------------------
    def has_field(self, key: str) -> bool:
        return bool(self._L.hfm_io_has_field(self._io, key.encode()))

    def erase_field(self, key: str) -> None:
        self._L.hfm_io_erase_field(self._io, key.encode())

    def get_string(self, key: str) -> str:
        n = self._L.hfm_io_get_string(self._io, key.encode(), None, 0)
        if n < 0:
            raise PyExp(self._err())
        buf = ctypes.create_string_buffer(n + 1)
        self._L.hfm_io_get_string(self._io, key.encode(), buf, n + 1)
        return buf.value.decode()

    def get_scalar(self, key: str) -> float:
        v = ctypes.c_double()
        if self._L.hfm_io_get_scalar(self._io, key.encode(), ctypes.byref(v)):
            raise PyExp(self._err())
        return v.value

    def get_array(self, key: str) -> np.ndarray:
        nd = ctypes.c_int()
        shape = (ctypes.c_int64 * 16)()
        if self._L.hfm_io_get_array_shape(self._io, key.encode(), ctypes.byref(nd), shape, 16):
            raise PyExp(self._err())
        dims = tuple(shape[i] for i in range(nd.value))
        out = np.empty(dims, dtype=np.float64)
        if self._L.hfm_io_get_array(self._io, key.encode(), _dptr(out), out.size):
            raise PyExp(self._err())
        return out

    def __getitem__(self, key):
        kind = self._L.hfm_io_field_kind(self._io, key.encode())
        if kind == 0:
            raise KeyError(key)
        if kind == 1:
            return self.get_string(key)
        a = self.get_array(key)
        return float(a) if a.ndim == 0 else a

    def computed_keys(self) -> list[str]:
        n = self._L.hfm_io_computed_keys(self._io, None, 0)
        buf = ctypes.create_string_buffer(max(n, 0) + 1)
        self._L.hfm_io_computed_keys(self._io, buf, len(buf))
        s = buf.value.decode()
        return s.split() if s else []

    def run(self) -> None:
        if self._L.hfm_run(self._io):
            raise PyExp(self._err())

    def messages(self) -> str:
        """Warnings and notes of the last run (what the reference writes to its log stream)."""
        return (self._L.hfm_io_messages(self._io) or b"").decode()


def run(params: dict) -> dict:
    """Solve an HFM input dictionary on the GPU; returns the computed keys (RowMajor default)."""
    io = HFMIO(default_ordering="RowMajor" if "arrayOrdering" not in params else "")
    for k, v in params.items():
        io[k] = v
    io.run()
    return {k: io[k] for k in io.computed_keys()}


# ---------------------------------------------------------------------------------------------
# Device-resident plan over torch tensors
# ---------------------------------------------------------------------------------------------
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise PyExp("no GPU visible: the hamiltonfastmarching_amd compute path runs on the GPU only")
    return torch


def _stream_ptr(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Plan:
    """Device-resident solver for one grid (C ABI hfm_plan_*).

    Inputs are torch CUDA float64 tensors (row-major like the reference's RowMajor ordering).
    """

    def __init__(self, model: str, dims, grid_scale, origin=None, precision: int = 32, xi: float = 1.0,
                 eps: float = 0.1):
        self._L = _lib.load()
        self.model = model
        self.dims = tuple(int(n) for n in dims)
        self.dim = MODEL_DIM[model]
        gs = np.broadcast_to(np.asarray(grid_scale, np.float64), (len(self.dims),)).copy()
        d = (ctypes.c_int64 * len(self.dims))(*self.dims)
        org = None if origin is None else np.ascontiguousarray(origin, np.float64)
        h = ctypes.c_void_p()
        rc = self._L.hfm_plan_create(ctypes.byref(h), model.encode(), d, _dptr(gs),
                                     None if org is None else _dptr(org), int(precision), float(xi), float(eps))
        if rc:
            raise PyExp(_lib.last_error())
        self._h = h
        self._keep = (gs, org)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.hfm_plan_destroy(h)
            self._h = None

    def _check(self, rc):
        if rc:
            raise PyExp(_lib.last_error())

    def set_metric(self, metric, dual: bool = False):
        torch = _torch()
        m = metric.to(device="cuda", dtype=torch.float64).contiguous()
        K = 3 if self.dim == 2 else 6
        is_field = int(m.numel() != K)
        self._check(self._L.hfm_plan_set_metric(self._h, ctypes.c_void_p(m.data_ptr()), is_field, int(dual),
                                                _stream_ptr(torch)))
        return self

    def set_cost(self, cost):
        torch = _torch()
        c = cost.to(device="cuda", dtype=torch.float64).contiguous()
        is_field = int(c.numel() != 1)
        self._check(self._L.hfm_plan_set_cost(self._h, ctypes.c_void_p(c.data_ptr()), is_field,
                                              _stream_ptr(torch)))
        return self

    def set_walls(self, walls, refine: bool = False):
        """Obstacles (walls key): a boolean/numeric field of the grid's shape (nonzero = wall), or
        None to remove them; refine = refineStencilAtWallBoundary."""
        torch = _torch()
        if walls is None:
            self._check(self._L.hfm_plan_set_walls(self._h, None, int(refine), _stream_ptr(torch)))
            self._walls = None
            return self
        w = torch.as_tensor(walls).to("cuda")
        if tuple(w.shape) != tuple(self.dims):
            raise PyExp("Error : Field walls has inconsistent dimensions.")
        w = (w != 0).to(torch.uint8).contiguous()
        self._check(self._L.hfm_plan_set_walls(self._h, ctypes.c_void_p(w.data_ptr()), int(refine),
                                               _stream_ptr(torch)))
        self._walls = w  # the copy is made on the stream; keep the source alive until then
        return self

    def solve(self, seeds, seed_values=None):
        torch = _torch()
        s = np.ascontiguousarray(np.asarray(seeds, np.float64).reshape(-1, self.dim))
        sv = None if seed_values is None else np.ascontiguousarray(seed_values, np.float64)
        self._check(self._L.hfm_plan_solve(self._h, len(s), _dptr(s), None if sv is None else _dptr(sv),
                                           _stream_ptr(torch)))
        return self

    def load_values(self, values, active):
        """Use cached arrival times and activeNeighs (grid-shaped) instead of a solve."""
        torch = _torch()
        v = torch.as_tensor(values, dtype=torch.float64).to("cuda").contiguous()
        a = torch.as_tensor(active, dtype=torch.float64).to("cuda").contiguous()
        if tuple(v.shape) != tuple(self.dims) or tuple(a.shape) != tuple(self.dims):
            raise PyExp("hfm_plan_load_values: values and activeNeighs must have the grid's shape")
        self._check(self._L.hfm_plan_load_values(self._h, ctypes.c_void_p(v.data_ptr()),
                                                 ctypes.c_void_p(a.data_ptr()), _stream_ptr(torch)))
        return self

    def values(self, out=None):
        torch = _torch()
        if out is None:
            out = torch.empty(self.dims, dtype=torch.float64, device="cuda")
        self._check(self._L.hfm_plan_values(self._h, ctypes.c_void_p(out.data_ptr()), _stream_ptr(torch)))
        return out

    def flow(self, want_flow=True, want_active=True):
        torch = _torch()
        fl = torch.empty(self.dims + (self.dim,), dtype=torch.float64, device="cuda") if want_flow else None
        ac = torch.empty(self.dims, dtype=torch.float64, device="cuda") if want_active else None
        self._check(self._L.hfm_plan_flow(self._h, ctypes.c_void_p(fl.data_ptr() if fl is not None else 0),
                                          ctypes.c_void_p(ac.data_ptr() if ac is not None else 0),
                                          _stream_ptr(torch)))
        return fl, ac

    def active_offsets(self):
        """activeOffsets (HFMInterface.hxx:554-577): dims + (nslots,) FP64 on the device."""
        torch = _torch()
        ns = self._L.hfm_plan_n_active_slots(self._h)
        out = torch.empty(self.dims + (ns,), dtype=torch.float64, device="cuda")
        self._check(self._L.hfm_plan_active_offsets(self._h, ctypes.c_void_p(out.data_ptr()), _stream_ptr(torch)))
        return out

    def residual(self):
        """Final FP64 residual check of the last solve: (largest |Lambda(u) - u| / |u| over the
        reached non-seed points, number of points where exactly one of Lambda(u), u is finite)."""
        torch = _torch()
        d = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        self._check(self._L.hfm_plan_residual(self._h, ctypes.byref(d), ctypes.byref(n), _stream_ptr(torch)))
        return d.value, n.value

    def geodesics(self, tips, step=0.25, threshold=0.001, volume_bound=0.0):
        """Geodesics backtracked from `tips` (ntips x dim physical points) by the discrete solver
        (GeodesicDiscreteSolver.hxx): (geodesicPoints (M x dim), geodesicLengths (ntips))."""
        torch = _torch()
        t = np.ascontiguousarray(np.asarray(tips, np.float64).reshape(-1, self.dim))
        lengths = np.zeros(len(t), np.int64)
        npts = ctypes.c_int64(0)
        nfail = ctypes.c_int(0)
        cap = 1 << 14
        while True:
            pts = np.zeros((cap, self.dim))
            self._check(self._L.hfm_plan_geodesics(self._h, len(t), _dptr(t), float(step), float(threshold),
                                                   float(volume_bound), _dptr(pts), cap,
                                                   lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                   ctypes.byref(npts), ctypes.byref(nfail), _stream_ptr(torch)))
            if npts.value <= cap:
                return pts[:npts.value].copy(), lengths.copy()
            cap = npts.value

    def geodesics_ode(self, tips, step=0.25, causality_tolerance=4.0):
        """Geodesics from `tips` by the ODE solver (GeodesicODESolver.hxx:156-217, midpoint steps
        along the interpolated flow): (geodesicPoints (M x dim), geodesicLengths (ntips),
        failed (ntips bools: the tips the reference reports as failed))."""
        torch = _torch()
        t = np.ascontiguousarray(np.asarray(tips, np.float64).reshape(-1, self.dim))
        lengths = np.zeros(len(t), np.int64)
        failed = np.zeros(len(t), np.int32)
        npts = ctypes.c_int64(0)
        cap = 1 << 14
        while True:
            pts = np.zeros((cap, self.dim))
            self._check(self._L.hfm_plan_geodesics_ode(
                self._h, len(t), _dptr(t), float(step), float(causality_tolerance), _dptr(pts), cap,
                lengths.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(npts),
                failed.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _stream_ptr(torch)))
            if npts.value <= cap:
                return pts[:npts.value].copy(), lengths.copy(), failed.astype(bool)
            cap = npts.value

    def adjoint(self, targets, weights=None, max_seeds=4096):
        torch = _torch()
        t = np.ascontiguousarray(np.asarray(targets, np.float64).reshape(-1, self.dim))
        w = np.ones(len(t)) if weights is None else np.ascontiguousarray(weights, np.float64)
        out = torch.empty(self.dims, dtype=torch.float64, device="cuda")
        si = np.zeros(max_seeds, np.int64)
        sw = np.zeros(max_seeds)
        ns = ctypes.c_int(0)
        self._check(self._L.hfm_plan_adjoint(self._h, len(t), _dptr(t), _dptr(w), ctypes.c_void_p(out.data_ptr()),
                                             si.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _dptr(sw),
                                             max_seeds, ctypes.byref(ns), _stream_ptr(torch)))
        n = min(ns.value, max_seeds)
        return out, (si[:n].copy(), sw[:n].copy())

    def forward_variation(self, cost_variation=None, seed_value_variation=None, nvar=None):
        """valueVariation (FirstVariation::ForwardVariation, FirstVariation.h:389-429) of the last
        solve: dims + (nvar,) FP64 on the device.  cost_variation: dims + (nvar,) (device tensor or
        array) or None; seed_value_variation: (nvar, nseeds) for the seeds of the last solve, or
        None."""
        torch = _torch()
        cv = None
        if cost_variation is not None:
            cv = torch.as_tensor(cost_variation, dtype=torch.float64).to("cuda").contiguous()
            if tuple(cv.shape[:-1]) != tuple(self.dims):
                raise PyExp("Forward variation error : inconsistent input sizes")
        svv = None
        if seed_value_variation is not None:
            svv = np.ascontiguousarray(np.asarray(seed_value_variation, np.float64))
            if svv.ndim != 2:
                raise PyExp("Error : Inconsistent size of seedValueVariation")
        nv = nvar
        for n in ([cv.shape[-1]] if cv is not None else []) + ([svv.shape[0]] if svv is not None else []):
            if nv is not None and n != nv:
                raise PyExp("Forward variation error : inconsistent input sizes")
            nv = n
        if nv is None:
            raise PyExp("Forward variation error : inconsistent input sizes")
        out = torch.empty(tuple(self.dims) + (nv,), dtype=torch.float64, device="cuda")
        self._check(self._L.hfm_plan_forward_variation(
            self._h, int(nv), ctypes.c_void_p(cv.data_ptr()) if cv is not None else None,
            int(svv.shape[1]) if svv is not None else 0, _dptr(svv) if svv is not None else None,
            ctypes.c_void_p(out.data_ptr()), _stream_ptr(torch)))
        return out

    def stats(self) -> dict:
        st = _lib.HfmPlanStats()
        self._check(self._L.hfm_plan_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def stencils(self):
        """(offsets int8 [N, K, d], weights float64 [N, K]) of a Riemann plan, row-major points."""
        N = int(np.prod(self.dims))
        K = 3 if self.dim == 2 else 6
        off = np.zeros((N, K, self.dim), np.int8)
        w = np.zeros((N, K))
        self._check(self._L.hfm_plan_get_stencils(self._h, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                                  _dptr(w)))
        return off, w

    def set_max_sweeps(self, sweeps: int):
        self._check(self._L.hfm_plan_set_max_sweeps(self._h, int(sweeps)))
        return self

    def set_delta(self, factor: float):
        self._check(self._L.hfm_plan_set_delta(self._h, float(factor)))
        return self

    def set_grid(self, blocks: int):
        self._check(self._L.hfm_plan_set_grid(self._h, int(blocks)))
        return self

    def set_round_target(self, tiles: int):
        self._check(self._L.hfm_plan_set_round_target(self._h, int(tiles)))
        return self

    def set_timing(self, enable: bool = True):
        self._check(self._L.hfm_plan_set_timing(self._h, int(enable)))
        return self

    def set_graph(self, enable: bool = True):
        """Run the update loop as a replayed HIP graph (default) or as individual launches."""
        self._check(self._L.hfm_plan_set_graph(self._h, int(enable)))
        return self

    def set_fused_schedule(self, enable=True):
        """Rounds list their own tiles inside the update kernel (True), or a separate listing
        kernel schedules each round (False); None restores the default (the first on 2-D
        grids, the second on 3-D grids)."""
        self._check(self._L.hfm_plan_set_fused_schedule(self._h, -1 if enable is None else int(bool(enable))))
        return self

    def set_order(self, order: int):
        """Scheme order (the `order` key): 1, or 2 / 3 for the recompute at acceptance with second /
        third order differences, solved by the reference's serial acceptance loop on the device
        (FP64 plans without walls)."""
        self._check(self._L.hfm_plan_set_order(self._h, int(order)))
        return self

    def set_factoring(self, values=None, gradients=None, mask=None, shift=None):
        """Static factor (factoringValues / factoringGradients / factoringMask /
        factoringIndexShift): values over a box (its shape), gradients (box + (dim,)) in grid-index
        units, mask (nonzero = on; None: everywhere), shift = the box's first grid index (None: the
        box is the whole grid).  values None removes the factor.  A factored solve runs the serial
        acceptance loop (FP64 plans without walls)."""
        if values is None:
            self._check(self._L.hfm_plan_set_factoring(self._h, None, None, None, None, None))
            return self
        v = np.ascontiguousarray(values, np.float64)
        g = np.ascontiguousarray(gradients, np.float64)
        if v.ndim != self.dim or g.shape != v.shape + (self.dim,):
            raise PyExp("set_factoring: values must be a box of the grid and gradients (box..., dim)")
        bd = np.array(v.shape, np.int64)
        sh = None if shift is None else np.ascontiguousarray(np.asarray(shift, np.int64).ravel())
        m = None if mask is None else np.ascontiguousarray(np.asarray(mask) != 0, np.uint8)
        if m is not None and m.shape != v.shape:
            raise PyExp("set_factoring: mask must have the shape of values")
        P = ctypes.POINTER
        self._check(self._L.hfm_plan_set_factoring(
            self._h, bd.ctypes.data_as(P(ctypes.c_int64)),
            None if sh is None else sh.ctypes.data_as(P(ctypes.c_int64)), _dptr(v), _dptr(g),
            None if m is None else m.ctypes.data_as(P(ctypes.c_uint8))))
        return self

    def set_stopping(self, max_accepted=-1, stop_distance=float("inf"), any_idx=(), all_idx=(),
                     any_unoriented=(), all_unoriented=()):
        """Stopping criteria (nMaxAccepted, stopAtDistance, stopWhenAnyAccepted / AllAccepted and
        their _Unoriented forms), as row-major grid indices ((x, y) indices x * n1 + y for the
        _unoriented lists of the curvature models).  No argument removes them.  A solve with a
        criterion runs the serial acceptance loop; stats()["n_accepted"] is the reference's
        nAccepted."""
        lists = [np.ascontiguousarray(np.asarray(a, np.int64).ravel())
                 for a in (any_idx, all_idx, any_unoriented, all_unoriented)]
        args = []
        for a in lists:
            args += [len(a), a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))]
        self._check(self._L.hfm_plan_set_stopping(self._h, int(max_accepted), float(stop_distance), *args))
        return self


class EikonalSolve:
    """``EikonalSolve.apply(cost, plan, seeds, targets, weights)`` -> sum_t weights[t] u(targets[t]).

    A ``torch.autograd.Function`` for the curvature models (cost field): forward runs the GPU
    solve, backward runs the adjoint kernel and returns d/d cost, and forward-mode AD (``jvp``)
    runs the forward-variation kernels with the cost tangent as costVariation
    (FirstVariation::ForwardVariation)."""

    _fn = None

    @classmethod
    def apply(cls, cost, plan, seeds, targets, weights=None):
        if cls._fn is None:
            cls._fn = _make_autograd()
        return cls._fn.apply(cost, plan, seeds, targets, weights)


def _make_autograd():
    import torch

    class _Eik(torch.autograd.Function):
        @staticmethod
        def forward(ctx, cost, plan, seeds, targets, weights):
            plan.set_cost(cost.detach())
            plan.solve(seeds)
            u = plan.values()
            t = np.asarray(targets, np.float64).reshape(-1, plan.dim)
            w = np.ones(len(t)) if weights is None else np.asarray(weights, np.float64)
            sens, _ = plan.adjoint(t, w)
            ctx.save_for_backward(sens)
            # value = sum_t w_t u(index(t)), index by the reference's seed/target snapping
            idx = _snap(plan, t)
            ctx.fwd = (plan, idx, w)
            val = torch.zeros((), dtype=torch.float64, device=u.device)
            for (i, wt) in zip(idx, w):
                val = val + wt * u[i]
            return val

        @staticmethod
        def backward(ctx, g):
            (sens,) = ctx.saved_tensors
            return g * sens, None, None, None, None

        @staticmethod
        def jvp(ctx, cost_t, plan_t, seeds_t, targets_t, weights_t):
            plan, idx, w = ctx.fwd
            vv = plan.forward_variation(cost_t.detach().to(torch.float64).contiguous()[..., None])[..., 0]
            out = torch.zeros((), dtype=torch.float64, device=vv.device)
            for (i, wt) in zip(idx, w):
                out = out + wt * vv[i]
            return out

    return _Eik


def _snap(plan, pts):
    """Grid index of physical points (ADim + floor, periodic theta), as for targets/seeds."""
    import math
    out = []
    st = plan._keep[0]
    org = plan._keep[1]
    for p in pts:
        idx = []
        for k in range(plan.dim):
            if plan.model in ("ReedsSheppForward2", "Elastica2") and k == 2:
                ds = 2 * math.pi / plan.dims[2]
                q = (p[k] + ds / 2) / ds
                idx.append(int(math.floor(q)) % plan.dims[2])
            else:
                o = 0.0 if org is None else org[k]
                idx.append(int(math.floor((p[k] - o) / st[k])))
        out.append(tuple(idx))
    return out

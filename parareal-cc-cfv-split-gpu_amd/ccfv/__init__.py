"""Python binding of the gfx950 T2S4 CC-CFV solver (C ABI: include/ccfv.h).

The binding is a thin ctypes layer over lib/libccfv.so.  It mirrors the
reference's solver entry points:

  GpuOneStep(ctx, tn, dt)             T2S4_Parareal.cu:785  -> Context.step
  CoarsePropagator(ctx, tn, DT)       T2S4_Parareal.cu:971  -> Context.coarse
  FinePropagator(ctx, tn, DT, J)      T2S4_Parareal.cu:977  -> Context.fine
  serial fine loop                    T2S4_Parareal.cu:1234 -> Context.plain_solve
  Parareal iterations                 T2S4_Parareal.cu:1250 -> Context.parareal_solve

Errors raise CcfvError carrying the C status code and message.  There is no
CPU fallback: without the built library (or without a GPU) every call fails.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import types
from .types import (IC, ParaReport, Problem, STATUS_NAMES, CCFV_OK,  # noqa: F401
                    CCFV_ENOTCONV)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get(
    "CCFV_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libccfv.so"))

_dp = ctypes.POINTER(ctypes.c_double)

# every exported symbol of include/ccfv.h, with its ctypes signature
_SIGS = {
    "ccfv_last_error": (ctypes.c_char_p, []),
    "ccfv_version": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)] * 2),
    "ccfv_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    "ccfv_problem_reference": (ctypes.c_int, [ctypes.POINTER(Problem),
                                              ctypes.c_int32, ctypes.c_double]),
    "ccfv_ic_reference": (ctypes.c_int, [ctypes.POINTER(IC), ctypes.c_double]),
    "ccfv_create": (ctypes.c_int, [ctypes.POINTER(Problem),
                                   ctypes.POINTER(ctypes.c_void_p)]),
    "ccfv_destroy": (None, [ctypes.c_void_p]),
    "ccfv_set_state": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "ccfv_get_state": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "ccfv_set_state_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ccfv_get_state_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "ccfv_state_device_ptr": (ctypes.c_int, [ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_void_p)]),
    "ccfv_fill_exact": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(IC),
                                       ctypes.c_double]),
    "ccfv_exact_to_host": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(IC),
                                          ctypes.c_double, _dp]),
    "ccfv_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                 ctypes.c_double]),
    "ccfv_coarse": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                   ctypes.c_double]),
    "ccfv_fine": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int32]),
    "ccfv_plain_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int32]),
    "ccfv_plain_solve_noalloc": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_void_p,
                                                ctypes.c_double, ctypes.c_double,
                                                ctypes.c_int32, ctypes.c_void_p]),
    "ccfv_advance": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int32, ctypes.c_int32]),
    "ccfv_parareal_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_double,
                                           ctypes.POINTER(ParaReport)]),
    "ccfv_parareal_correction": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4
                                 + [ctypes.c_size_t]),
    "ccfv_parareal_iterate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _dp]),
    "ccfv_sync": (ctypes.c_int, [ctypes.c_void_p]),
    "ccfv_time_sweep": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]),
    "ccfv_dispatch_floor": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_double)]),
    "ccfv_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64]),
    "ccfv_get_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                       ctypes.POINTER(ctypes.c_int64)]),
    "ccfv_get_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "ccfv_set_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "ccfv_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                     ctypes.POINTER(ctypes.c_int32)]),
    "ccfv_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "ccfv_profile_read": (ctypes.c_int, [ctypes.c_void_p,
                                         ctypes.POINTER(types.KernelStat)]),
    "ccfv_batch_create": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "ccfv_batch_destroy": (None, [ctypes.c_void_p]),
    "ccfv_batch_size": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "ccfv_batch_context": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "ccfv_batch_plain": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp,
                                        ctypes.POINTER(ctypes.c_int32)]),
    "ccfv_batch_parareal": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp,
                                           ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32),
                                           ctypes.c_int32, ctypes.c_double,
                                           ctypes.POINTER(ParaReport)]),
    "ccfv_batch_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                             ctypes.c_int64]),
    "ccfv_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp,
                                           ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
    "ccfv_request_bytes": (ctypes.c_int, [ctypes.POINTER(types.Request),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "ccfv_server_create": (ctypes.c_int, [ctypes.POINTER(types.ServerConfig),
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "ccfv_server_destroy": (None, [ctypes.c_void_p]),
    "ccfv_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(types.Request),
                                          _dp, ctypes.POINTER(ctypes.c_int64)]),
    "ccfv_server_step": (ctypes.c_int, [ctypes.c_void_p,
                                        ctypes.POINTER(types.ServerStatus)]),
    "ccfv_server_next_done": (ctypes.c_int, [ctypes.c_void_p,
                                             ctypes.POINTER(ctypes.c_int64)]),
    "ccfv_server_result": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _dp,
                                          ctypes.POINTER(ParaReport),
                                          ctypes.POINTER(ctypes.c_int32)]),
}

_lib = None


class CcfvError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def load(path: str | None = None):
    """Load lib/libccfv.so (raises if it is not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libccfv.so not found at {p}: build it with "
            "`make -C parareal-cc-cfv-split-gpu_amd/csrc` (or __graft_entry__.build())")
    L = ctypes.CDLL(p)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = L
    return L


def exported_symbols():
    return list(_SIGS)


def _check(rc: int, what: str):
    if rc != CCFV_OK:
        msg = load().ccfv_last_error().decode(errors="replace")
        raise CcfvError(rc, f"{what}: {msg}")


def device_count() -> int:
    n = ctypes.c_int32(0)
    rc = load().ccfv_device_count(ctypes.byref(n))
    return n.value if rc == CCFV_OK else 0


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def plain_solve_noalloc(problem: Problem, u_dev: int, t0: float, t1: float, nt: int,
                        stream: int = 0):
    """ccfv_plain_solve_noalloc on a device field at address u_dev (nx*ny
    doubles), on a HIP stream (0: the default stream)."""
    _check(load().ccfv_plain_solve_noalloc(ctypes.byref(problem), ctypes.c_void_p(u_dev),
                                           t0, t1, nt, ctypes.c_void_p(stream or None)),
           "ccfv_plain_solve_noalloc")


class Context:
    """One solver context on the current HIP device (ccfv_create)."""

    def __init__(self, problem: Problem):
        self.problem = problem.copy()
        h = ctypes.c_void_p()
        _check(load().ccfv_create(ctypes.byref(self.problem), ctypes.byref(h)),
               "ccfv_create")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            load().ccfv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.problem.nx, self.problem.ny)

    def set_state(self, u: np.ndarray):
        a = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        if a.size != self.problem.nx * self.problem.ny:
            raise ValueError("state has the wrong size")
        _check(load().ccfv_set_state(self._h, _ptr(a)), "ccfv_set_state")

    def get_state(self) -> np.ndarray:
        out = np.empty(self.problem.nx * self.problem.ny, dtype=np.float64)
        _check(load().ccfv_get_state(self._h, _ptr(out)), "ccfv_get_state")
        return out

    def state_ptr(self) -> int:
        """Device address of the state (nx*ny doubles)."""
        v = ctypes.c_void_p()
        _check(load().ccfv_state_device_ptr(self._h, ctypes.byref(v)),
               "ccfv_state_device_ptr")
        return int(v.value)

    def set_state_device(self, ptr: int):
        _check(load().ccfv_set_state_device(self._h, ctypes.c_void_p(ptr)),
               "ccfv_set_state_device")

    def get_state_device(self, ptr: int):
        _check(load().ccfv_get_state_device(self._h, ctypes.c_void_p(ptr)),
               "ccfv_get_state_device")

    def fill_exact(self, ic: IC, t: float = 0.0):
        _check(load().ccfv_fill_exact(self._h, ctypes.byref(ic), t),
               "ccfv_fill_exact")

    def exact(self, ic: IC, t: float) -> np.ndarray:
        out = np.empty(self.problem.nx * self.problem.ny, dtype=np.float64)
        _check(load().ccfv_exact_to_host(self._h, ctypes.byref(ic), t, _ptr(out)),
               "ccfv_exact_to_host")
        return out

    def step(self, tn: float, h: float):
        _check(load().ccfv_step(self._h, tn, h), "ccfv_step")

    def coarse(self, tn: float, DT: float):
        _check(load().ccfv_coarse(self._h, tn, DT), "ccfv_coarse")

    def fine(self, tn: float, DT: float, J: int):
        _check(load().ccfv_fine(self._h, tn, DT, J), "ccfv_fine")

    def plain_solve(self, t0: float, t1: float, nt: int):
        _check(load().ccfv_plain_solve(self._h, t0, t1, nt), "ccfv_plain_solve")

    def advance(self, t_origin: float, dt: float, n_begin: int, n_end: int):
        _check(load().ccfv_advance(self._h, t_origin, dt, n_begin, n_end), "ccfv_advance")

    def parareal_solve(self, t0, t1, nt_fine, n_slices, kmax, tol=0.0,
                       raise_on_notconv=False) -> ParaReport:
        rep = ParaReport()
        rc = load().ccfv_parareal_solve(self._h, t0, t1, nt_fine, n_slices, kmax,
                                        tol, ctypes.byref(rep))
        if rc == CCFV_ENOTCONV and not raise_on_notconv:
            return rep
        _check(rc, "ccfv_parareal_solve")
        return rep

    def parareal_iterate(self, k: int) -> np.ndarray:
        """U^k(t1) of iteration k of the last parareal_solve (needs
        OPT_KEEP_ITERATES on)."""
        out = np.empty(self.problem.nx * self.problem.ny, dtype=np.float64)
        _check(load().ccfv_parareal_iterate(self._h, k, _ptr(out)), "ccfv_parareal_iterate")
        return out

    def parareal_correction(self, g_new: int, g_old: int, f: int, u: int, n: int):
        """u = g_new + (f - g_old) over n doubles at device addresses, queued
        on the context's stream (ccfv_parareal_correction)."""
        _check(load().ccfv_parareal_correction(
            self._h, ctypes.c_void_p(g_new), ctypes.c_void_p(g_old), ctypes.c_void_p(f),
            ctypes.c_void_p(u), ctypes.c_size_t(n)), "ccfv_parareal_correction")

    def sync(self):
        _check(load().ccfv_sync(self._h), "ccfv_sync")

    def time_sweep(self, direction: int, fused: bool, nfields: int, h: float,
                   launches: int, layout: int = 0, tile=(0, 0), graph: bool = False) -> float:
        """µs per launch of back-to-back sweeps of one class (ccfv_time_sweep)."""
        us = ctypes.c_double()
        _check(load().ccfv_time_sweep(self._h, direction, int(fused), nfields, h, launches,
                                      layout, tile[0], tile[1], int(graph), ctypes.byref(us)),
               "ccfv_time_sweep")
        return us.value

    def dispatch_floor(self, direction: int, kind: int, launches: int,
                       graph: bool = False) -> float:
        """µs per launch of back-to-back empty (kind 0) or copy (kind 1)
        launches with the one-field sweep's grid, or of y and x copies
        alternating in place (kind 2, a plain solve's pattern without its
        arithmetic) (ccfv_dispatch_floor)."""
        us = ctypes.c_double()
        _check(load().ccfv_dispatch_floor(self._h, direction, kind, launches, int(graph),
                                          ctypes.byref(us)), "ccfv_dispatch_floor")
        return us.value

    def set_option(self, key: int, value: int):
        _check(load().ccfv_set_option(self._h, key, value), "ccfv_set_option")

    def get_option(self, key: int) -> int:
        v = ctypes.c_int64()
        _check(load().ccfv_get_option(self._h, key, ctypes.byref(v)), "ccfv_get_option")
        return v.value

    def get_tiles(self) -> list:
        t = (ctypes.c_int32 * 8)()
        _check(load().ccfv_get_tiles(self._h, t), "ccfv_get_tiles")
        return list(t)

    def set_tiles(self, tiles):
        t = (ctypes.c_int32 * 8)(*tiles)
        _check(load().ccfv_set_tiles(self._h, t), "ccfv_set_tiles")

    def autotune(self, batch_fields: int, h: float) -> list:
        """Time every valid launch shape and keep the fastest (see ccfv.h)."""
        t = (ctypes.c_int32 * 8)()
        _check(load().ccfv_autotune(self._h, batch_fields, h, t), "ccfv_autotune")
        return list(t)

    def profile_enable(self, on: bool = True):
        _check(load().ccfv_profile_enable(self._h, int(on)), "ccfv_profile_enable")

    def profile_read(self) -> dict:
        arr = (types.KernelStat * len(types.KSTAT_NAMES))()
        _check(load().ccfv_profile_read(self._h, arr), "ccfv_profile_read")
        return {types.KSTAT_NAMES[s.kind]: {"launches": s.launches,
                                            "total_ms": s.total_ms,
                                            "bytes": s.bytes} for s in arr}


class _MemberContext(Context):
    """A context owned by a Batch: usable like a Context, never destroyed
    from Python (ccfv_batch_destroy frees it)."""

    def __init__(self, handle, problem: Problem, owner):
        self.problem = problem.copy()
        self._h = handle
        self._owner = owner  # keeps the batch alive while this view exists

    def close(self):
        self._h = None


def _per_problem(v, n, dtype):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=dtype), (n,)))
    return a


class Batch:
    """Independent problems of mixed sizes solved together (ccfv_batch_*):
    every sub-step of every problem goes through shared ragged launches, and
    each problem's result is bitwise the one its own Context would give."""

    def __init__(self, problems):
        arr = (Problem * len(problems))(*[p.copy() for p in problems])
        h = ctypes.c_void_p()
        _check(load().ccfv_batch_create(arr, len(problems), ctypes.byref(h)),
               "ccfv_batch_create")
        self._h = h
        self.problems = [p.copy() for p in problems]
        self.contexts = []
        for i, p in enumerate(self.problems):
            c = ctypes.c_void_p()
            _check(load().ccfv_batch_context(self._h, i, ctypes.byref(c)),
                   "ccfv_batch_context")
            self.contexts.append(_MemberContext(c, p, self))

    def __len__(self):
        return len(self.problems)

    def __getitem__(self, i) -> Context:
        return self.contexts[i]

    def close(self):
        if self._h:
            for c in self.contexts:
                c.close()
            load().ccfv_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def size(self) -> int:
        n = ctypes.c_int32()
        _check(load().ccfv_batch_size(self._h, ctypes.byref(n)), "ccfv_batch_size")
        return n.value

    def set_option(self, key: int, value: int):
        """ccfv_set_option on every member (e.g. OPT_SERIAL_FINISH)."""
        _check(load().ccfv_batch_set_option(self._h, key, value), "ccfv_batch_set_option")

    def autotune(self, t0, t1, nt, batch_fields: int = 16):
        """Launch shapes per line class (ccfv_batch_autotune)."""
        n = len(self)
        a0 = _per_problem(t0, n, np.float64)
        a1 = _per_problem(t1, n, np.float64)
        an = _per_problem(nt, n, np.int32)
        _check(load().ccfv_batch_autotune(self._h, _ptr(a0), _ptr(a1),
                                          an.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          batch_fields), "ccfv_batch_autotune")

    def plain(self, t0, t1, nt):
        n = len(self)
        a0 = _per_problem(t0, n, np.float64)
        a1 = _per_problem(t1, n, np.float64)
        an = _per_problem(nt, n, np.int32)
        _check(load().ccfv_batch_plain(self._h, _ptr(a0), _ptr(a1),
                                       an.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))),
               "ccfv_batch_plain")

    def parareal(self, t0, t1, nt_fine, n_slices, kmax, tol=0.0,
                 raise_on_notconv=False) -> list:
        n = len(self)
        a0 = _per_problem(t0, n, np.float64)
        a1 = _per_problem(t1, n, np.float64)
        af = _per_problem(nt_fine, n, np.int32)
        asl = _per_problem(n_slices, n, np.int32)
        reps = (ParaReport * n)()
        i32 = ctypes.POINTER(ctypes.c_int32)
        rc = load().ccfv_batch_parareal(self._h, _ptr(a0), _ptr(a1),
                                        af.ctypes.data_as(i32), asl.ctypes.data_as(i32),
                                        kmax, tol, reps)
        if not (rc == CCFV_ENOTCONV and not raise_on_notconv):
            _check(rc, "ccfv_batch_parareal")
        return list(reps)


def request_bytes(req: "types.Request") -> int:
    """Device bytes a request holds while resident in a Server."""
    b = ctypes.c_uint64()
    _check(load().ccfv_request_bytes(ctypes.byref(req), ctypes.byref(b)),
           "ccfv_request_bytes")
    return b.value


class Server:
    """The serving engine (ccfv_server_*): continuous batching of Parareal
    solve requests under a device-memory budget."""

    def __init__(self, mem_budget: int = 0, max_active: int = 0,
                 max_slices: int = 0, max_bypass: int = 0, adaptive_above: int = 0,
                 serial_finish: int = 0):
        cfg = types.ServerConfig(mem_budget, max_active, max_slices, max_bypass,
                                 adaptive_above, serial_finish)
        h = ctypes.c_void_p()
        _check(load().ccfv_server_create(ctypes.byref(cfg), ctypes.byref(h)),
               "ccfv_server_create")
        self._h = h
        self._shape = {}

    def close(self):
        if self._h:
            load().ccfv_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, req: "types.Request", u0: np.ndarray | None = None) -> int:
        rid = ctypes.c_int64()
        p = None
        if u0 is not None:
            a = np.ascontiguousarray(u0, dtype=np.float64).reshape(-1)
            if a.size != req.problem.nx * req.problem.ny:
                raise ValueError("u0 has the wrong size")
            p = _ptr(a)
        _check(load().ccfv_server_submit(self._h, ctypes.byref(req), p,
                                         ctypes.byref(rid)), "ccfv_server_submit")
        self._shape[rid.value] = req.problem.nx * req.problem.ny
        return rid.value

    def step(self) -> "types.ServerStatus":
        st = types.ServerStatus()
        _check(load().ccfv_server_step(self._h, ctypes.byref(st)), "ccfv_server_step")
        return st

    def next_done(self) -> int:
        rid = ctypes.c_int64()
        _check(load().ccfv_server_next_done(self._h, ctypes.byref(rid)),
               "ccfv_server_next_done")
        return rid.value

    def result(self, rid: int):
        """(state, report, status) of a finished request; releases it."""
        out = np.empty(self._shape[rid], dtype=np.float64)
        rep = ParaReport()
        st = ctypes.c_int32()
        _check(load().ccfv_server_result(self._h, rid, _ptr(out), ctypes.byref(rep),
                                         ctypes.byref(st)), "ccfv_server_result")
        del self._shape[rid]
        return out, rep, st.value


__all__ = ["Context", "Batch", "Server", "CcfvError", "load", "device_count",
           "types", "Problem", "IC", "ParaReport", "exported_symbols",
           "request_bytes", "plain_solve_noalloc"]

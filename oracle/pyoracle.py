"""ctypes wrapper of the CPU oracle (oracle/libccfv_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PKG = os.path.join(os.path.dirname(_HERE), "parareal-cc-cfv-split-gpu_amd")
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)

from ccfv.types import IC, Problem  # noqa: E402

LIB_PATH = os.path.join(_HERE, "libccfv_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Problem)
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_intc.restype = ctypes.c_double
        L.orc_intc.argtypes = [ctypes.c_double] * 4
        L.orc_ppm_coef.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_ghosts.argtypes = [_dp, ctypes.c_int, _dp]
        L.orc_deriv.restype = ctypes.c_double
        L.orc_deriv.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_thomas.argtypes = [ctypes.c_double] * 3 + [_dp, ctypes.c_int, _dp]
        L.orc_cyclic.argtypes = [ctypes.c_double] * 2 + [_dp, ctypes.c_int, _dp]
        L.orc_line_sweep.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int, _dp]
        L.orc_feet_x.argtypes = [P, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, _dp]
        L.orc_feet_y.argtypes = L.orc_feet_x.argtypes
        L.orc_sweep_x.argtypes = [P, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int]
        L.orc_sweep_y.argtypes = [P, _dp, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_step.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double]
        L.orc_plain.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int]
        L.orc_parareal.argtypes = [P, _dp, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_double, _dp,
                                   ctypes.POINTER(ctypes.c_int), _dp]
        L.orc_exact.argtypes = [P, ctypes.POINTER(IC), ctypes.c_double, _dp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def intc(c6, dltc, cl, xi) -> float:
    return lib().orc_intc(c6, dltc, cl, xi)


def ppm_coef(line, i, periodic=False):
    line = np.ascontiguousarray(line, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_ppm_coef(_ptr(line), len(line), int(periodic), int(i), _ptr(out))
    return out


def ghosts(line):
    line = np.ascontiguousarray(line, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_ghosts(_ptr(line), len(line), _ptr(out))
    return out


def deriv(line, iL, xi, d, periodic=False, quirk=False) -> float:
    line = np.ascontiguousarray(line, dtype=np.float64)
    return lib().orc_deriv(_ptr(line), len(line), int(periodic), int(quirk),
                           int(iL), xi, d)


def thomas(a, b, c, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    x = np.zeros_like(r)
    lib().orc_thomas(a, b, c, _ptr(r), len(r), _ptr(x))
    return x


def cyclic(a, b, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    x = np.zeros_like(r)
    lib().orc_cyclic(a, b, _ptr(r), len(r), _ptr(x))
    return x


def line_sweep(c, feet, lo, hi, *, periodic, quirk=False, eps=1e-15, nu, h,
               theta, deriv_postremap=False):
    c = np.ascontiguousarray(c, dtype=np.float64)
    feet = np.ascontiguousarray(feet, dtype=np.float64)
    out = np.zeros_like(c)
    rc = lib().orc_line_sweep(_ptr(c), _ptr(feet), len(c), lo, hi,
                              int(periodic), int(quirk), eps, nu, h, theta,
                              int(deriv_postremap), _ptr(out))
    _check(rc, "line_sweep")
    return out


def feet_x(p: Problem, j, tau, t_hi, t_lo):
    out = np.zeros(p.nx + 1)
    lib().orc_feet_x(ctypes.byref(p), j, tau, t_hi, t_lo, _ptr(out))
    return out


def feet_y(p: Problem, i, tau, t_hi, t_lo):
    out = np.zeros(p.ny + 1)
    lib().orc_feet_y(ctypes.byref(p), i, tau, t_hi, t_lo, _ptr(out))
    return out


def exact(p: Problem, ic: IC, t: float) -> np.ndarray:
    out = np.zeros(p.nx * p.ny)
    _check(lib().orc_exact(ctypes.byref(p), ctypes.byref(ic), t, _ptr(out)),
           "exact")
    return out


def sweep_x(p: Problem, c, tn, h, substep=1):
    c = np.ascontiguousarray(c, dtype=np.float64)
    out = np.zeros_like(c)
    _check(lib().orc_sweep_x(ctypes.byref(p), _ptr(c), _ptr(out), tn, h,
                             substep), "sweep_x")
    return out


def sweep_y(p: Problem, c, tn, h):
    c = np.ascontiguousarray(c, dtype=np.float64)
    out = np.zeros_like(c)
    _check(lib().orc_sweep_y(ctypes.byref(p), _ptr(c), _ptr(out), tn, h),
           "sweep_y")
    return out


def step(p: Problem, c, tn, h):
    c = np.array(c, dtype=np.float64, copy=True)
    _check(lib().orc_step(ctypes.byref(p), _ptr(c), tn, h), "step")
    return c


def plain(p: Problem, c, t0, t1, nt):
    c = np.array(c, dtype=np.float64, copy=True)
    _check(lib().orc_plain(ctypes.byref(p), _ptr(c), t0, t1, nt), "plain")
    return c


def parareal(p: Problem, c, t0, t1, nt_fine, n_slices, kmax, tol=0.0,
             want_slices=False):
    c = np.array(c, dtype=np.float64, copy=True)
    incr = np.zeros(max(kmax, 1))
    iters = ctypes.c_int(0)
    sl = np.zeros(p.nx * p.ny * (n_slices + 1)) if want_slices else None
    rc = lib().orc_parareal(ctypes.byref(p), _ptr(c), t0, t1, nt_fine, n_slices,
                            kmax, tol, _ptr(incr), ctypes.byref(iters),
                            _ptr(sl) if sl is not None else None)
    _check(rc, "parareal")
    res = {"u": c, "increments": incr[: iters.value].tolist(),
           "iterations": iters.value}
    if sl is not None:
        res["slices"] = sl.reshape(n_slices + 1, p.nx * p.ny)
    return res

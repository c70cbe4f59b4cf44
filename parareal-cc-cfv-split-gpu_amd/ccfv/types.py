"""ctypes mirrors of the C ABI structs in include/ccfv.h and the problem presets.

The presets restate the reference problem (T2S4_Parareal.cu:27-61, the config
keys of :1031-1049) and the five benchmark workloads on periodic [0,1]^2.
"""
from __future__ import annotations

import ctypes
import math

# status codes (include/ccfv.h)
CCFV_OK = 0
CCFV_EINVAL = -1
CCFV_ENOMEM = -2
CCFV_EDEVICE = -3
CCFV_ENOTCONV = -4
CCFV_ECFL = -5
CCFV_EBUSY = -6
CCFV_ENONFINITE = -7

STATUS_NAMES = {
    CCFV_OK: "OK", CCFV_EINVAL: "EINVAL", CCFV_ENOMEM: "ENOMEM",
    CCFV_EDEVICE: "EDEVICE", CCFV_ENOTCONV: "ENOTCONV", CCFV_ECFL: "ECFL",
    CCFV_EBUSY: "EBUSY", CCFV_ENONFINITE: "ENONFINITE",
}

BC_REFERENCE = 0
BC_PERIODIC = 1

VEL_SOLID_BODY = 0
VEL_CONSTANT = 1
VEL_SWIRL = 2

FLAG_STEP3_POSTREMAP = 0x1
FLAG_DERIV_GHOST_QUIRK = 0x2
FLAGS_REFERENCE = FLAG_STEP3_POSTREMAP | FLAG_DERIV_GHOST_QUIRK

SRC_NONE = 0
SRC_CONSTANT = 1
SRC_SINSIN = 2

IC_ROT_GAUSSIAN = 0
IC_GAUSSIANS = 1
IC_SINSIN = 2
IC_COSINE_BELL = 3

MAX_ITERS = 64

OPT_PIPELINE = 1
OPT_PIPE_GROUP = 2
OPT_GRAPHS = 3
OPT_PLAIN_GRAPHS = 4
OPT_FUSE = 5
OPT_PANEL = 6
OPT_SERIAL_FINISH = 7
OPT_KEEP_ITERATES = 8


class Problem(ctypes.Structure):
    """ccfv_problem."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("bc", ctypes.c_int32), ("vel_kind", ctypes.c_int32),
        ("x_lo", ctypes.c_double), ("x_hi", ctypes.c_double),
        ("y_lo", ctypes.c_double), ("y_hi", ctypes.c_double),
        ("nu", ctypes.c_double), ("vel", ctypes.c_double * 4),
        ("flags", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
        ("src_kind", ctypes.c_int32), ("src_reserved", ctypes.c_uint32),
        ("src", ctypes.c_double * 5),
    ]

    def copy(self, **kw) -> "Problem":
        p = Problem()
        ctypes.pointer(p)[0] = self
        for k, v in kw.items():
            if k in ("vel", "src"):
                arr = getattr(p, k)
                for i in range(len(arr)):
                    arr[i] = 0.0
                for i, x in enumerate(v):
                    arr[i] = x
            else:
                setattr(p, k, v)
        return p

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny

    @property
    def dx(self) -> float:
        return (self.x_hi - self.x_lo) / self.nx

    @property
    def dy(self) -> float:
        return (self.y_hi - self.y_lo) / self.ny

    def __repr__(self) -> str:
        return (f"Problem(nx={self.nx}, ny={self.ny}, bc={self.bc}, "
                f"vel_kind={self.vel_kind}, box=[{self.x_lo},{self.x_hi}]x"
                f"[{self.y_lo},{self.y_hi}], nu={self.nu}, "
                f"vel={list(self.vel)}, flags={self.flags})")


class IC(ctypes.Structure):
    """ccfv_ic."""
    _fields_ = [("kind", ctypes.c_int32), ("n", ctypes.c_int32),
                ("p", ctypes.c_double * 16)]

    def __repr__(self) -> str:
        return f"IC(kind={self.kind}, n={self.n}, p={list(self.p)})"


class ParaReport(ctypes.Structure):
    """ccfv_parareal_report."""
    _fields_ = [
        ("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
        ("n_slices", ctypes.c_int32), ("j_fine", ctypes.c_int32),
        ("increment", ctypes.c_double * MAX_ITERS),
        ("fine_steps_run", ctypes.c_double),
        ("coarse_steps_run", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("serial_finish", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]

    @property
    def increments(self) -> list:
        return [self.increment[i] for i in range(self.iterations)]


class ServerConfig(ctypes.Structure):
    """ccfv_server_config."""
    _fields_ = [("mem_budget", ctypes.c_uint64), ("max_active", ctypes.c_int32),
                ("max_slices", ctypes.c_int32), ("max_bypass", ctypes.c_int32),
                ("adaptive_above", ctypes.c_int32), ("serial_finish", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class Request(ctypes.Structure):
    """ccfv_request: one Parareal solve for the serving engine."""
    _fields_ = [("problem", Problem), ("ic", IC), ("t0", ctypes.c_double),
                ("t1", ctypes.c_double), ("nt_fine", ctypes.c_int32),
                ("n_slices", ctypes.c_int32), ("kmax", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("tol", ctypes.c_double)]


class ServerStatus(ctypes.Structure):
    """ccfv_server_status."""
    _fields_ = [("active", ctypes.c_int32), ("queued", ctypes.c_int32),
                ("finished", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("mem_used", ctypes.c_uint64), ("mem_budget", ctypes.c_uint64),
                ("steps", ctypes.c_uint64), ("last_step_ms", ctypes.c_double)]


def make_request(problem: Problem, ic: IC, t0: float, t1: float, nt_fine: int,
                 n_slices: int, kmax: int, tol: float = 0.0) -> Request:
    r = Request()
    r.problem = problem.copy()
    r.ic = ic
    r.t0, r.t1 = t0, t1
    r.nt_fine, r.n_slices, r.kmax = nt_fine, n_slices, kmax
    r.tol = tol
    return r


KSTAT_NAMES = ["x_batch", "y_batch", "x_single", "y_single", "x_epilog"]


class KernelStat(ctypes.Structure):
    """ccfv_kernel_stat."""
    _fields_ = [("kind", ctypes.c_int32), ("launches", ctypes.c_int32),
                ("total_ms", ctypes.c_double), ("bytes", ctypes.c_double)]


def make_problem(nx, ny, *, bc, vel_kind, vel, nu, box=(0.0, 1.0, 0.0, 1.0),
                 flags=0) -> Problem:
    p = Problem()
    p.nx, p.ny, p.bc, p.vel_kind = nx, ny, bc, vel_kind
    p.x_lo, p.x_hi, p.y_lo, p.y_hi = box
    p.nu = nu
    for i, v in enumerate(vel):
        p.vel[i] = v
    p.flags = flags
    return p


def make_ic(kind, params, n=0) -> IC:
    ic = IC()
    ic.kind, ic.n = kind, n
    for i, v in enumerate(params):
        ic.p[i] = v
    return ic


# --------------------------------------------------------------------------
# the reference problem (T2S4_Parareal.cu:27-33, :44-56, config keys N and K)
# --------------------------------------------------------------------------
def reference_problem(n: int, K: float) -> Problem:
    return make_problem(n, n, bc=BC_REFERENCE, vel_kind=VEL_SOLID_BODY,
                        vel=(4.0, 0.0, 0.0), nu=K, box=(-1.0, 1.0, -1.0, 1.0),
                        flags=FLAGS_REFERENCE)


def reference_ic() -> IC:
    # C_Exact at t=0: Gaussian exp(-r^2/sigma), x0=-0.35, y0=0, sigma=0.005
    return make_ic(IC_ROT_GAUSSIAN, (-0.35, 0.0, 0.005, 4.0, 0.0, 0.0))


# --------------------------------------------------------------------------
# benchmark workloads (periodic [0,1]^2)
# --------------------------------------------------------------------------
def advection_sinsin_problem(n: int, *, flags: int = 0) -> Problem:
    """Constant velocity (1, 0.5), nu=0.01, u0 = sin(2 pi x) sin(2 pi y)."""
    return make_problem(n, n, bc=BC_PERIODIC, vel_kind=VEL_CONSTANT,
                        vel=(1.0, 0.5), nu=0.01, flags=flags)


def advection_sinsin_ic() -> IC:
    return make_ic(IC_SINSIN, (1.0, 0.5, 0.01))


def advection_sinsin_steps(n: int, T: float = 1.0, cfl: float = 0.8) -> int:
    """Fine steps at CFL 0.8 against max |u| = 1: dt = 0.8 dx."""
    return int(round(T / (cfl * (1.0 / n))))


def rotation_problem(n: int, *, nu: float = 1e-3, flags: int = 0) -> Problem:
    """Solid-body rotation u=-2 pi (y-1/2), v=2 pi (x-1/2)."""
    return make_problem(n, n, bc=BC_PERIODIC, vel_kind=VEL_SOLID_BODY,
                        vel=(2.0 * math.pi, 0.5, 0.5), nu=nu, flags=flags)


def rotation_ic(x0=0.5, y0=0.75, width=0.05) -> IC:
    """Gaussian hump, standard deviation `width`: exp(-r^2/(2 w^2))."""
    return make_ic(IC_ROT_GAUSSIAN,
                   (x0, y0, 2.0 * width * width, 2.0 * math.pi, 0.5, 0.5))


def swirl_problem(n: int, *, T: float = 5.0, nu: float = 1e-4,
                  flags: int = 0) -> Problem:
    return make_problem(n, n, bc=BC_PERIODIC, vel_kind=VEL_SWIRL, vel=(T,),
                        nu=nu, flags=flags)


def cosine_bell_ic(x0=0.5, y0=0.75, R=0.15) -> IC:
    return make_ic(IC_COSINE_BELL, (x0, y0, R))


def constant_gaussians(nx, ny, u, v, nu, bumps, flags=0):
    """Constant-velocity problem with a sum of Gaussians (x0, y0, w, a)."""
    p = make_problem(nx, ny, bc=BC_PERIODIC, vel_kind=VEL_CONSTANT,
                     vel=(u, v), nu=nu, flags=flags)
    params = [u, v, nu]
    for b in bumps:
        params.extend(b)
    return p, make_ic(IC_GAUSSIANS, params, n=len(bumps))


def manufactured_sinsin(n: int, *, u: float = 1.0, v: float = 0.5, nu: float = 0.01,
                        A: float = 0.5, w: float = 2.0 * math.pi, flags: int = 0):
    """Constant velocity (u, v), diffusion nu, and a source whose exact
    solution is the decaying translated sin*sin plus A sin(w t) S(x) S(y)
    (S(z) = sin(2 pi z)).  The scheme adds 2 h f per step (every sub-step
    adds theta h (fbar0 + f1)), so f is half of the equation's source
    g = A [w cos(w t) + 8 pi^2 nu sin(w t)] S S
        + 2 pi A sin(w t) (u C(x) S(y) + v S(x) C(y)).
    Returns (problem, ic, exact_extra) with exact_extra(t) the amplitude
    A sin(w t) of the S(x) S(y) part."""
    p = advection_sinsin_problem(n, flags=flags).copy(vel=(u, v), nu=nu)
    tp = 2.0 * math.pi
    p.src_kind = SRC_SINSIN
    half = 0.5
    for i, x in enumerate((w, half * A * w, half * A * 2.0 * tp * tp * nu,
                           half * A * tp * u, half * A * tp * v)):
        p.src[i] = x
    ic = make_ic(IC_SINSIN, (u, v, nu))
    return p, ic, (lambda t: A * math.sin(w * t))

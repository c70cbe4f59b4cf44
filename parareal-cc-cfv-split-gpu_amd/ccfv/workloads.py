"""Seeded workload generators: the ragged batch of independent solves and
the request stream of the serving loop.

Ragged batch (many independent problems of uneven size): grid Nx, Ny each
uniform in {64, 128, 256, 512, 1024}; constant velocity components uniform in
[-1, 1]; nu log-uniform in [1e-4, 1e-2]; initial data a sum of 3 Gaussians
(centres uniform in the unit square, standard deviations uniform in
[0.03, 0.1], amplitudes uniform in [0.5, 1]); periodic [0,1]^2, FP64;
T = 0.5; Parareal slices uniform in {8, 16, 32}.  The fine step keeps the
CFL number at most 0.8 in both directions with at least J_MIN fine steps per
slice, so the coarse step is J fine steps.

Serving stream: Poisson arrivals of such solves, plus solid-body rotation
cases (the Parareal benchmark problem at 256^2..1024^2, shorter horizon).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import types as T

GRIDS = (64, 128, 256, 512, 1024)
SLICES = (8, 16, 32)
T_END = 0.5
FINE_CFL = 0.8
J_MIN = 4
RAGGED_SEED = 20261203


@dataclass
class Solve:
    """One Parareal solve: problem, initial data at t0, plan."""
    problem: T.Problem
    ic: T.IC
    t0: float
    t1: float
    nt_fine: int
    n_slices: int
    kind: str = "ragged"

    @property
    def cells(self) -> int:
        return self.problem.nx * self.problem.ny

    @property
    def cell_updates(self) -> int:
        """serial-equivalent work: cells x fine steps"""
        return self.cells * self.nt_fine

    def request(self, kmax: int, tol: float) -> T.Request:
        return T.make_request(self.problem, self.ic, self.t0, self.t1, self.nt_fine,
                              self.n_slices, kmax, tol)


def fine_steps(nx: int, ny: int, u: float, v: float, t_end: float, ns: int,
               cfl: float = FINE_CFL, j_min: int = J_MIN) -> int:
    """nt_fine = ns * J with CFL = |u| dt nx <= cfl (and likewise in y)."""
    need = t_end * max(abs(u) * nx, abs(v) * ny) / cfl
    J = max(j_min, int(math.ceil(need / ns)))
    return ns * J


def ragged_solve(rng: np.random.Generator, grids=GRIDS, slices=SLICES,
                 t_end=T_END) -> Solve:
    nx = int(rng.choice(grids))
    ny = int(rng.choice(grids))
    u, v = (float(x) for x in rng.uniform(-1.0, 1.0, 2))
    nu = float(10.0 ** rng.uniform(-4.0, -2.0))
    bumps = []
    for _ in range(3):
        x0, y0 = (float(x) for x in rng.uniform(0.0, 1.0, 2))
        w = float(rng.uniform(0.03, 0.1))
        a = float(rng.uniform(0.5, 1.0))
        bumps.append((x0, y0, w, a))
    ns = int(rng.choice(slices))
    p, ic = T.constant_gaussians(nx, ny, u, v, nu, bumps)
    return Solve(p, ic, 0.0, t_end, fine_steps(nx, ny, u, v, t_end, ns), ns)


def ragged_batch(count: int = 256, seed: int = RAGGED_SEED, **kw) -> list:
    rng = np.random.default_rng(seed)
    return [ragged_solve(rng, **kw) for _ in range(count)]


def rotation_solve(n: int, t_end: float = 0.125, slices: int = 16) -> Solve:
    """The Parareal benchmark problem at n^2 over [0, t_end]: dt = 1/1024
    (x-CFL pi n/1024 at the hump), J = nt / slices."""
    nt = int(round(t_end * 1024))
    nt = max(slices, nt // slices * slices)
    return Solve(T.rotation_problem(n), T.rotation_ic(), 0.0, t_end, nt, slices,
                 kind="rotation")


def serving_stream(count: int, rate: float, seed: int = RAGGED_SEED + 1,
                   rotation_share: float = 0.1, grids=GRIDS, slices=SLICES) -> list:
    """(arrival time, Solve) pairs: Poisson arrivals at `rate` per second;
    a `rotation_share` of them rotating-flow cases at 256^2..1024^2."""
    rng = np.random.default_rng(seed)
    times = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for t in times:
        if rng.uniform() < rotation_share:
            n = int(rng.choice((256, 512, 1024)))
            out.append((float(t), rotation_solve(n)))
        else:
            out.append((float(t), ragged_solve(rng, grids=grids, slices=slices)))
    return out

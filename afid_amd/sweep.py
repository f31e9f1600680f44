"""Batched parameter sweeps: many independent cells advanced together.

The reference runs one cell per job (main.F90).  Here every cell keeps its own
stream and its own captured step graph (afid_steps_async), and
afid_batch_steps queues step n of every cell before step n+1 of any, so the
kernels of small cells overlap on the GPU instead of running one cell after
another.  Each cell's dt comes from its own CFL number on the device, as in
main.F90:169-189.

The sweep distribution: nxm uniform in {32, 48, 64, 96, 128}, ny = nz uniform in
{64, 128, 192, 256}, Ra = 10^U(5, 7), Pr uniform in {0.7, 1, 4.4}, aspect ratio
ylen = zlen = 2, clipped-Chebyshev stretching istr3 = 6, str3 = 6, RK3,
CFL-limited dt (cflmax 1.2, dtmin 1e-8, dtmax 1e-2), cell c seeded with
100 + c (conduction profile plus +-1e-3 noise).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from .engine import Cell, CellParams, batch_steps

SWEEP_NXM = (32, 48, 64, 96, 128)
SWEEP_NYZ = (64, 128, 192, 256)
SWEEP_PR = (0.7, 1.0, 4.4)
# main.F90 step control used by sweeps and serving (bou.in: IDTV=1, CFLMAX=1.2,
# DTMIN=1e-8, DTMAX=1e-2)
STEP_CTL = dict(idtv=1, dt=1e-3, dtmin=1e-8, dtmax=1e-2, limitcfl=1.2)


@dataclass
class CellSpec:
    params: CellParams
    seed: int
    steps: int

    @property
    def cell_updates(self) -> int:
        return self.params.cells * self.steps


def sweep_specs(n: int = 64, steps: int = 500, seed: int = 0) -> List[CellSpec]:
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n):
        nyz = int(rng.choice(SWEEP_NYZ))
        p = CellParams(nxm=int(rng.choice(SWEEP_NXM)), nym=nyz, nzm=nyz, nsst=3, istr3=6, str3=6.0, alx3=1.0,
                       ylen=2.0, zlen=2.0, ray=float(10 ** rng.uniform(5, 7)), pra=float(rng.choice(SWEEP_PR)))
        out.append(CellSpec(p, 100 + c, steps))
    return out


@dataclass
class SweepResult:
    mode: str
    cells: int
    seconds: float
    cell_updates: int
    steps_wanted: List[int] = field(default_factory=list)
    errorcodes: List[int] = field(default_factory=list)
    steps_done: List[int] = field(default_factory=list)
    max_divergence: float = 0.0

    @property
    def cells_per_hour(self) -> float:
        return self.cells / self.seconds * 3600.0

    @property
    def mcups(self) -> float:
        return self.cell_updates / self.seconds / 1e6

    def as_dict(self) -> dict:
        return {"mode": self.mode, "cells": self.cells, "seconds": round(self.seconds, 3),
                "cells_per_hour": round(self.cells_per_hour, 1), "Mcell_updates_per_s": round(self.mcups, 2),
                "max_divergence": self.max_divergence, "errorcodes": sorted(set(self.errorcodes)),
                "all_steps_done": self.steps_done == self.steps_wanted}


def run_batched(specs: List[CellSpec], on_done: Optional[Callable[[int, Cell], None]] = None,
                tune: bool = True) -> SweepResult:
    """All cells resident at once; steps queued interleaved across the cells.
    tune: each cell's stencil launch shape is picked by afid_tune first."""
    t0 = time.perf_counter()
    cells = [Cell(s.params) for s in specs]
    try:
        ctl = STEP_CTL
        for s, c in zip(specs, cells):
            if tune:
                c.tune()
            c.init_linear_noise(s.seed, 1e-3)
            c.steps_async(0, 0, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
        # equal step counts (the sweep's case) go in one interleaved call;
        # the remainder of longer cells follows
        base = min(s.steps for s in specs)
        batch_steps(cells, base)
        for s, c in zip(specs, cells):
            if s.steps > base:
                c.steps_more(s.steps - base)
        res = SweepResult("batched", len(specs), 0.0, sum(s.cell_updates for s in specs), [s.steps for s in specs])
        for q, c in enumerate(cells):
            err, _, done = c.steps_wait()
            res.errorcodes.append(err)
            res.steps_done.append(done)
            res.max_divergence = max(res.max_divergence, c.max_divergence())
            if on_done:
                on_done(q, c)
    finally:
        for c in cells:
            c.close()
    res.seconds = time.perf_counter() - t0
    return res


def run_serial(specs: List[CellSpec], host_driven: bool = True) -> SweepResult:
    """One cell after another.  host_driven: the plain drop-in path (afid_step per
    step, the CFL number read back every step, as the Fortran host runs it);
    otherwise one afid_steps call per cell (captured step graphs)."""
    t0 = time.perf_counter()
    ctl = STEP_CTL
    res = SweepResult("serial_host" if host_driven else "serial_graph", len(specs), 0.0,
                      sum(s.cell_updates for s in specs), [s.steps for s in specs])
    for s in specs:
        c = Cell(s.params)
        try:
            c.init_linear_noise(s.seed, 1e-3)
            if host_driven:
                err, done = 0, 0
                for n in range(s.steps):
                    err, _ = c.step(n, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
                    done += 1
                    if err:
                        break
            else:
                err, _, done = c.steps(0, s.steps, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"],
                                       ctl["limitcfl"])
            res.errorcodes.append(err)
            res.steps_done.append(done)
            res.max_divergence = max(res.max_divergence, c.max_divergence())
        finally:
            c.close()
    res.seconds = time.perf_counter() - t0
    return res

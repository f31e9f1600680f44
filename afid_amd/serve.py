"""Serving: simulations admitted as they arrive, by device-memory budget.

A request is one Rayleigh-Benard cell (parameters, seed, number of main.F90
steps).  The loop admits waiting requests first-come first-served while the
estimated device bytes of the admitted cells stay within the budget
(afid_estimate_bytes), keeps every admitted cell resident, and keeps each one's
stream fed with chunks of captured steps (afid_steps_async / afid_steps_more),
so that work of cells admitted at different times overlaps on the GPU.  A cell
finishes when all its steps have run, or as soon as its step control has
stopped with an error code, as main.F90 does; its memory then goes back to the
budget and the next waiting request is admitted.  A library error raised for
one cell (a refused factorisation, say) ends that request only.  Each cell's result depends only on its own
request: the cells share no device data.

The serving stream: Poisson arrivals at lam requests/s; nxm uniform in
[32, 192]; ny = nz uniform over the 2-3-5-smooth sizes in [64, 384]; Ra =
10^U(5, 8); Pr uniform in [0.7, 7]; steps uniform in [200, 2000]; aspect ratio
2, istr3 = 6, str3 = 6, CFL-limited dt.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np

from .engine import AfidError, Cell, CellParams, estimate_bytes
from .sweep import STEP_CTL, CellSpec


def smooth_sizes(lo: int, hi: int) -> List[int]:
    """n in [lo, hi] with no prime factor above 5 (the FFT sizes, factorize.F90)."""
    out = []
    for n in range(lo, hi + 1):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    spec: CellSpec


def serving_requests(n: int, lam: float = 4.0, seed: int = 0, steps_range=(200, 2000)) -> List[Request]:
    rng = np.random.default_rng(seed)
    sizes = smooth_sizes(64, 384)
    t = 0.0
    out = []
    for r in range(n):
        t += rng.exponential(1.0 / lam)
        nyz = int(rng.choice(sizes))
        p = CellParams(nxm=int(rng.integers(32, 193)), nym=nyz, nzm=nyz, nsst=3, istr3=6, str3=6.0, alx3=1.0,
                       ylen=2.0, zlen=2.0, ray=float(10 ** rng.uniform(5, 8)), pra=float(rng.uniform(0.7, 7.0)))
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        out.append(Request(r, t, CellSpec(p, 1000 + r, steps)))
    return out


@dataclass
class Completion:
    rid: int
    arrival: float
    admitted: float
    finished: float
    steps_done: int
    errorcode: int
    max_divergence: float
    cell_updates: int
    refused: str = ""       # why the library refused the cell (the reference would abort the job)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Active:
    req: Request
    cell: object
    queued: int
    admitted: float
    nbytes: int


class Server:
    """Continuous-batching request loop over resident cells.

    make_cell(params) -> object with tune, init_linear_noise, steps_async,
    steps_more, busy, steps_wait, max_divergence, close (engine.Cell by default);
    estimate(params) -> bytes; on_finish(request, cell) runs before the cell is
    closed (the tests read the fields there)."""

    def __init__(self, budget_bytes: float = 200e9, chunk: int = 50,
                 make_cell: Callable = Cell, estimate: Callable = estimate_bytes,
                 clock: Callable[[], float] = time.perf_counter, idle_sleep: float = 2e-4,
                 on_finish: Optional[Callable[[Request, object], None]] = None, tune: bool = True):
        if chunk < 1:
            raise ValueError("chunk must be at least 1 step")
        self.budget = budget_bytes
        self.chunk = chunk
        self.make_cell = make_cell
        self.estimate = estimate
        self.clock = clock
        self.idle_sleep = idle_sleep
        self.on_finish = on_finish
        self.tune = tune                # each admitted cell's stencil launch shape from afid_tune
        self.peak_bytes = 0
        self.peak_active = 0

    def run(self, requests: List[Request]) -> List[Completion]:
        for r in requests:
            if self.estimate(r.spec.params) > self.budget:
                raise ValueError(f"request {r.rid} needs more device memory than the whole budget")
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque = deque()
        active: Dict[int, _Active] = {}
        used = 0
        done: List[Completion] = []
        ctl = STEP_CTL
        t0 = self.clock()
        while pending or waiting or active:
            now = self.clock() - t0
            progressed = False
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # first come, first served: the head waits until it fits
            while waiting and used + self.estimate(waiting[0].spec.params) <= self.budget:
                r = waiting.popleft()
                nb = self.estimate(r.spec.params)
                progressed = True
                cell = None
                try:
                    cell = self.make_cell(r.spec.params)
                    if self.tune:
                        cell.tune()
                    cell.init_linear_noise(r.spec.seed, 1e-3)
                    first = min(self.chunk, r.spec.steps)
                    cell.steps_async(0, first, ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"],
                                     ctl["limitcfl"])
                except AfidError as e:
                    # e.g. the singular zero pressure mode, where the reference aborts
                    # (SolvePressureCorrection.F90:130-134): the request ends unrun
                    if cell is not None:
                        cell.close()
                    t = self.clock() - t0
                    done.append(Completion(r.rid, r.arrival, t, t, 0, e.code, 0.0, 0, str(e)))
                    continue
                active[r.rid] = _Active(r, cell, first, self.clock() - t0, nb)
                used += nb
            self.peak_bytes = max(self.peak_bytes, used)
            self.peak_active = max(self.peak_active, len(active))
            for rid in list(active):
                a = active[rid]
                try:
                    if a.cell.busy():
                        continue
                    progressed = True
                    # the stream is idle: reading the step state costs one small copy
                    err, _, ndone = a.cell.steps_wait()
                    if err == 0 and a.queued < a.req.spec.steps:
                        # later chunks are twice the first: fewer idle gaps between polls
                        n = min(2 * self.chunk, a.req.spec.steps - a.queued)
                        a.cell.steps_more(n)
                        a.queued += n
                        continue
                    # all steps ran, or the step control halted with an error code
                    # (main.F90 stops there): the request ends now
                    div = a.cell.max_divergence()   # inf if a field is not finite
                    if self.on_finish:
                        self.on_finish(a.req, a.cell)
                    fin = Completion(rid, a.req.arrival, a.admitted, self.clock() - t0, ndone, err, div,
                                     a.req.spec.params.cells * ndone)
                except AfidError as e:
                    # e.g. AFID_NUM_PIVOT from a step: this request fails, the others go on
                    t = self.clock() - t0
                    fin = Completion(rid, a.req.arrival, a.admitted, t, 0, e.code, 0.0, 0, str(e))
                a.cell.close()
                used -= a.nbytes
                del active[rid]
                done.append(fin)
            if not progressed:
                if pending and not waiting and not active:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                else:
                    time.sleep(self.idle_sleep)
        return done


def summarize(done: List[Completion], makespan: Optional[float] = None) -> dict:
    """Latency percentiles over the requests that ran; refused ones are counted apart."""
    ran = [c for c in done if not c.refused]
    lat = np.array([c.latency for c in ran])
    span = makespan if makespan is not None else max(c.finished for c in done)
    updates = sum(c.cell_updates for c in ran)
    return {"requests": len(done), "refused": len(done) - len(ran),
            "p50_latency_s": float(np.percentile(lat, 50)) if len(ran) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(ran) else None, "makespan_s": span,
            "Mcell_updates_per_s": updates / span / 1e6,
            "max_divergence": max((c.max_divergence for c in ran), default=0.0),
            "errorcodes": sorted({c.errorcode for c in ran})}

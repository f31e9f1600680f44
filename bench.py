"""Throughput of AFiD's time advance on one MI355X.

Workload (default): production single cell, nxm=384 (clipped-Chebyshev
stretching istr3=6, str3=6), ny=nz=768 periodic, ylen=zlen=1, Ra=1e8, Pr=1,
RK3, CFL-limited dt (cflmax 1.2, dtmin 1e-8, dtmax 1e-2), seeded conduction
profile plus +-1e-3 noise, FP64.  One step = one main.F90 time step: CalcMaxCFL,
dt selection, three substeps.  The timed steps run on the device from one
captured graph per step (afid_steps); inputs are resident in HBM before timing.

Prints one JSON line (metric Mcell-updates/s = nxm*nym*nzm*steps/s/1e6) with a
roofline block for the dominant kernel (HIP events on the engine's stream), a
CPU baseline from the reference's own routines on a bounded sample of the same
workload, and beside it the C restatement on one full step of the whole cell.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic FP64 words per cell of each kernel: the distinct fields it must
# read and write once (8 B each).  stencil_rhs reads vx,vy,vz,T,pr and the four
# histories and writes the histories and the four right-hand sides b + A q_old;
# column_solve reads those and writes the four new fields (u* in place).
# Spectral arrays hold nym/2+1 complex
# values per y line, one word per cell.  The correction reads u*, dph and pr and
# writes the velocities and pr in place.  The substep figure is the reference's
# per-routine compulsory traffic: explicit VX 5, VY 4, VZ 4, Temp 5; implicit
# VX/VY/VZ 6 each, Temp 5; divergence 4; pressure solve 10 (5 passes);
# CorrectVelocity 7; CorrectPressure 3.
WORDS = {
    # RK3: the first substep (ro = 0) does not read the four histories (13
    # words); the third does not store them, since the next step's first
    # substep does not read them (13); the second does both (17).  The
    # HIP-event average covers all three.
    "stencil_rhs": (13 + 17 + 13) / 3,
    "column_solve": 4 + 4,
    "div_yfft_r2c": 3 + 1,
    "zfft_fwd": 2,
    "xsolve_pressure": 2,
    "zfft_bwd": 2,
    "yfft_c2r": 2,
    "correct_vel_pr": 5 + 4,
}
SUBSTEP_WORDS = 65

# Measured HBM bytes per launch (rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by a
# 1 GiB copy; profiles/substep_kernels_cfg2_384x768x768_pmc.md), per bench
# kernel name: the PMC kernels it stands for and their weights per substep.
PMC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_cfg2_384x768x768.json")
PMC_KERNELS = {
    # k_stencil_rhs<offset type, histories read, device scalars, FP32 histories>: the timed
    # graph steps run the device-scalar instances
    "stencil_rhs": [("k_stencil_rhs<u32, F, T, F>", 1 / 3), ("k_stencil_rhs<u32, T, T, F>", 2 / 3)],
    "column_solve": [("k_column_solve<25, u32>", 1.0)],
    "correct_vel_pr": [("k_correct<F>", 1.0)],
    "div_yfft_r2c": [("k_yfft_forward<double>", 1.0)],
    "zfft_fwd": [("k_zfft<double>", 1.0)],
    "zfft_bwd": [("k_zfft<double>", 1.0)],
    "xsolve_pressure": [("k_xsolve_pressure<32, 12, double>", 1.0)],
    "yfft_c2r": [("k_yfft_inverse<double>", 1.0)],
}


def measured_traffic(config: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed PMC summary (cfg2 only)."""
    if config != "cfg2" or kernel not in PMC_KERNELS or not os.path.exists(PMC_FILE):
        return None
    with open(PMC_FILE) as f:
        k = json.load(f)["kernels"]
    try:
        return round(sum(w * k[name]["traffic_bytes"] for name, w in PMC_KERNELS[kernel]))
    except KeyError:
        return None

CONFIGS = {
    "cfg2": dict(nxm=384, nym=768, nzm=768, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e8, pra=1.0, seed=2),
    "cfg1": dict(nxm=64, nym=128, nzm=128, istr3=6, str3=6.0, ylen=2.0, zlen=2.0, ray=1e6, pra=1.0, seed=1),
}


REF_SAMPLE_NYZ = 128     # lateral size of the reference-routine sample (same x grid as the cell)
REF_SAMPLE_STEPS = 2


def cpu_baseline_reference(cfg: dict) -> dict | None:
    """The reference's own routines (oracle/_ref/ref_harness, built from the
    reference sources: TimeMarcher, ExplicitTerms*, ImplicitAndUpdate*,
    SolveImpEqnUpdate_* with LAPACK, corrections, CalcMaxCFL and the main.F90
    dt rule) on a bounded sample of the workload: the cell's x grid, Ra and Pr
    with ny = nz = REF_SAMPLE_NYZ, REF_SAMPLE_STEPS main.F90 steps from the same
    seeded state.  One MPI rank on one core: the harness stands in the C
    restatement's Poisson solve for the FFTW one (FFTW3 is not in the image),
    which needs the whole domain on one rank.  The step loop's own wall time is
    used (written by the harness), not process start-up or file I/O."""
    import tempfile
    from oracle import oracle as O
    if not O.ref_available():
        return None
    p = {k: v for k, v in cfg.items() if k != "seed"}
    p.update(nym=REF_SAMPLE_NYZ, nzm=REF_SAMPLE_NYZ, nsst=3, alx3=1.0, inslws=1, inslwn=1)
    with tempfile.TemporaryDirectory() as wd:
        o = O.Oracle(**p)
        o.init_linear_noise(cfg["seed"], 1e-3)
        with open(os.path.join(wd, "fields_in.bin"), "wb") as f:
            for n in ("vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp", "rhs"):
                f.write(o.arr(n).tobytes())
        o.close()
        O.run_ref(wd, p, mode=3, nsteps=REF_SAMPLE_STEPS, dt=1e-3, idtv=1, dtmin=1e-8, dtmax=1e-2,
                  limitcfl=1.2, timeout=600)
        with open(os.path.join(wd, "march_time.txt")) as f:
            t = float(f.read().split()[0])
    cells = p["nxm"] * p["nym"] * p["nzm"]
    return {"value": round(cells * REF_SAMPLE_STEPS / t / 1e6, 3), "unit": "Mcell-updates/s", "cores": 1,
            "kind": "reference",
            "sample": f"the reference's routines (amdflang -O2, MKL LAPACK, 1 MPI rank) on "
                      f"{p['nxm']}x{p['nym']}x{p['nzm']} (the cell's x grid, lateral size cut to "
                      f"{REF_SAMPLE_NYZ}), {REF_SAMPLE_STEPS} main.F90 steps from the seeded state, step loop "
                      f"{t:.1f} s; the Poisson solve is the C restatement's (FFTW3 absent)"}


def cpu_baseline_port(cfg: dict, threads: int) -> dict:
    """The C restatement (oracle, OpenMP over z planes) on one full main.F90 step
    (CalcMaxCFL, dt, all three RK3 substeps) of the whole cell."""
    from oracle import oracle as O
    O.lib().afo_set_threads(threads)
    p = {k: v for k, v in cfg.items() if k != "seed"}
    o = O.Oracle(nsst=3, alx3=1.0, **p)
    o.init_linear_noise(cfg["seed"], 1e-3)
    t0 = time.perf_counter()
    o.step(0, 1, 1e-3, 1e-8, 1e-2, 1.2)
    t = time.perf_counter() - t0
    cells = p["nxm"] * p["nym"] * p["nzm"]
    o.close()
    return {"value": round(cells / t / 1e6, 3), "unit": "Mcell-updates/s", "cores": threads,
            "kind": "port", "host_cpus": os.cpu_count(),
            "sample": f"one full RK3 step (3 substeps) of the same {p['nxm']}x{p['nym']}x{p['nzm']} cell, "
                      f"{t:.1f} s on {threads} OpenMP threads"}


def copy_calibration(mib: int) -> None:
    """One device-to-device copy of `mib` MiB (read + write of known size), for
    scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings of the engine's kernels."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(n))
    hip.hipDeviceSynchronize()
    hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def bench_sweep(args):
    """Batched sweep: all cells resident, steps interleaved; beside it the same
    cells one at a time through the plain host-driven path (afid_step)."""
    from afid_amd.sweep import run_batched, run_serial, sweep_specs
    specs = sweep_specs(args.sweep_cells, args.sweep_steps, seed=0)
    print(f"sweep: {len(specs)} cells x {args.sweep_steps} steps, batched", file=sys.stderr, flush=True)
    b = run_batched(specs)
    print("sweep: serial host-driven", file=sys.stderr, flush=True)
    h = run_serial(specs, host_driven=True)
    out = {
        "metric": "cells/hour", "value": round(b.cells_per_hour, 1), "unit": "cells/hour", "n_gpus": 1,
        "steps": args.sweep_steps, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded conduction + noise)",
        "config": {"workload": f"batched sweep: {len(specs)} cells, nxm in {{32..128}}, ny=nz in {{64..256}}, "
                               f"Ra 1e5-1e7, Pr in {{0.7,1,4.4}}, aspect 2, istr3=6 str3=6, CFL dt, "
                               f"{args.sweep_steps} steps each"},
        "batched": b.as_dict(), "serial_host": h.as_dict(),
        "speedup_vs_serial_host": round(h.seconds / b.seconds, 3),
        "max_divergence": max(b.max_divergence, h.max_divergence),
    }
    print(json.dumps(out))


def bench_mixed(args):
    """Mixed-precision cell against the FP64 run of the same seeded input:
    nxm=256 (istr3=6, str3=6), ny=nz=512, aspect 2, Ra=1e7, Pr=0.7, CFL dt;
    throughput of both and the per-field rel-L2 difference after the steps."""
    import numpy as np
    from afid_amd import Cell, CellParams
    from afid_amd.engine import PREC_FP64, PREC_MIXED
    p = dict(nxm=256, nym=512, nzm=512, istr3=6, str3=6.0, ylen=2.0, zlen=2.0, ray=1e7, pra=0.7)
    cells = p["nxm"] * p["nym"] * p["nzm"]
    ctl = (1, 1e-3, 1e-8, 1e-2, 1.2)
    res, fields = {}, {}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        print(f"mixed: {name} run, {args.mixed_steps} steps", file=sys.stderr, flush=True)
        c = Cell(CellParams(nsst=3, alx3=1.0, precision=prec, **p))
        c.init_linear_noise(3, 1e-3)
        c.steps(0, 2, *ctl)             # captures the step graph; dt = dtmin for these two
        c.synchronize()
        t0 = time.perf_counter()
        err, _, done = c.steps(2, args.mixed_steps - 2, *ctl)
        c.synchronize()
        el = time.perf_counter() - t0
        if err or done != args.mixed_steps - 2:
            raise RuntimeError(f"{name}: step control stopped: errorcode {err} after {done} steps")
        fields[name] = {nm: c.download(nm) for nm in ("vx", "vy", "vz", "temp", "pr")}
        res[name] = {"Mcell_updates_per_s": round(cells * done / el / 1e6, 2), "ms_per_step": round(el / done * 1e3, 3),
                     "max_divergence": c.CheckDivergence(), "time": c.time}
        # per-kernel HIP-event times over a few graph steps after the comparison
        c.profile(True)
        c.profile_reset()
        c.steps(args.mixed_steps, args.profile_steps, *ctl)
        prof = c.profile_read()
        c.profile(False)
        res[name]["kernel_ms_per_step"] = {k: round(v[0] / args.profile_steps, 3) for k, v in prof.items() if v[1]}
        c.close()

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    nxm = p["nxm"]
    err = {nm: rel(fields["mixed"][nm], fields["fp64"][nm]) for nm in ("vx", "vy", "vz", "temp")}
    pm, pf = fields["mixed"]["pr"][:, :, :nxm], fields["fp64"]["pr"][:, :, :nxm]
    err["pr"] = rel(pm - pm.mean(), pf - pf.mean())
    out = {
        "metric": "Mcell-updates/s", "value": res["mixed"]["Mcell_updates_per_s"], "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": args.mixed_steps, "warmup": 2, "ms_per_step": res["mixed"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32",
        "data": "synthetic (seeded conduction + noise)",
        "config": {"workload": f"mixed precision: RB cell {p['nxm']}x{p['nym']}x{p['nzm']} istr3=6 str3=6 aspect 2 "
                               f"Ra=1e7 Pr=0.7 RK3 CFL dt, {args.mixed_steps} steps; FP32 histories and FFT "
                               "spectrum, one refinement sweep per step", "cells": cells},
        "fp64": res["fp64"], "mixed": res["mixed"],
        "speedup_vs_fp64": round(res["mixed"]["Mcell_updates_per_s"] / res["fp64"]["Mcell_updates_per_s"], 3),
        "rel_l2_vs_fp64": err,
    }
    print(json.dumps(out))


def bench_serve(args):
    """Serving: Poisson arrivals (4/s) admitted by a 200 GB device budget."""
    from afid_amd.serve import Server, serving_requests, summarize
    reqs = serving_requests(args.requests, lam=4.0, seed=0)
    print(f"serve: {len(reqs)} requests", file=sys.stderr, flush=True)
    srv = Server(budget_bytes=200e9, chunk=50)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    span = time.perf_counter() - t0
    s = summarize(done, span)
    out = {
        "metric": "Mcell-updates/s", "value": round(s["Mcell_updates_per_s"], 2), "unit": "Mcell-updates/s",
        "n_gpus": 1, "steps": sum(r.spec.steps for r in reqs), "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded conduction + noise)",
        "config": {"workload": f"serving: {len(reqs)} requests, Poisson 4/s, nxm U[32,192], ny=nz 2-3-5 smooth "
                               f"in [64,384], Ra 1e5-1e8, Pr U[0.7,7], steps U[200,2000], 200 GB budget"},
        "p50_latency_s": round(s["p50_latency_s"], 3), "p99_latency_s": round(s["p99_latency_s"], 3),
        "refused": s["refused"],
        "makespan_s": round(span, 3), "peak_active": srv.peak_active, "peak_bytes": srv.peak_bytes,
        "max_divergence": s["max_divergence"], "errorcodes": s["errorcodes"],
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="run one device-to-device copy of MIB MiB first (PMC traffic calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="cfg2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    # the host's CPU share: OMP_NUM_THREADS where set (16 per GPU on the GPU hosts),
    # else every CPU this process may run on
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0)))
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--workload", default="cell", choices=("cell", "sweep", "serve", "mixed"),
                    help="cell: one production cell (default); sweep: 64-cell batched parameter sweep; "
                         "serve: Poisson request stream; mixed: mixed-precision cell against FP64")
    ap.add_argument("--mixed-steps", type=int, default=200)
    ap.add_argument("--sweep-cells", type=int, default=64)
    ap.add_argument("--sweep-steps", type=int, default=500)
    ap.add_argument("--requests", type=int, default=24)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(each cell runs on one device and is not sharded across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "sweep":
        return bench_sweep(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "mixed":
        return bench_mixed(args)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    from afid_amd import Cell, CellParams
    cfg = CONFIGS[args.config]
    p = {k: v for k, v in cfg.items() if k != "seed"}
    cell = Cell(CellParams(nsst=3, alx3=1.0, **p))
    cell.init_linear_noise(cfg["seed"], 1e-3)
    cells = p["nxm"] * p["nym"] * p["nzm"]
    ctl = dict(idtv=1, dt=1e-3, dtmin=1e-8, dtmax=1e-2, limitcfl=1.2)
    args_ctl = (ctl["idtv"], ctl["dt"], ctl["dtmin"], ctl["dtmax"], ctl["limitcfl"])
    # whole steps on the device: CFL, dt, factorisations and the three substeps
    # replayed from one captured graph per step; host sync only around the loop
    ntime = 0
    err, _, done = cell.steps(ntime, args.warmup, *args_ctl)
    ntime += done
    cell.synchronize()
    t0 = time.perf_counter()
    err, _, done = cell.steps(ntime, args.steps, *args_ctl)
    cell.synchronize()
    elapsed = time.perf_counter() - t0
    ntime += done
    if err or done != args.steps:
        raise RuntimeError(f"step control stopped early: errorcode {err} after {done} steps")
    ms_per_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed / 1e6
    divmax = cell.CheckDivergence()

    # per-kernel HIP-event timing on the engine stream (separate short pass):
    # steps from a second capture of the same step graph, each kernel
    # bracketed by event-record nodes, read after every replay
    cell.profile(True)
    cell.profile_reset()
    err, _, done = cell.steps(ntime, args.profile_steps, *args_ctl)
    if err or done != args.profile_steps:
        raise RuntimeError(f"profiled steps stopped early: errorcode {err} after {done} steps")
    ntime += done
    prof = cell.profile_read()
    cell.profile(False)
    kern = {k: {"avg_ms": v[0] / v[1], "launches": v[1]} for k, v in prof.items() if v[1] > 0}
    substep_ms = sum(v[0] for k, v in prof.items() if k in WORDS) / (3 * args.profile_steps)
    dom = max((k for k in kern if k in WORDS), key=lambda k: kern[k]["avg_ms"] * kern[k]["launches"])
    alg_bytes = round(WORDS[dom] * 8 * cells)
    achieved = alg_bytes / (kern[dom]["avg_ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args.config, dom),
                "kernel": dom, "alg_bytes_per_launch": alg_bytes, "avg_ms": round(kern[dom]["avg_ms"], 4)}
    substep_alg = SUBSTEP_WORDS * 8 * cells / (substep_ms * 1e-3) / 1e9
    cell.close()

    cpu = cpu_port = None
    if not args.no_cpu_baseline:
        print("cpu baseline: reference routines on a bounded sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline_reference(cfg)
        print(f"cpu baseline: C restatement on {args.cpu_threads} threads, one full step", file=sys.stderr,
              flush=True)
        cpu_port = cpu_baseline_port(cfg, args.cpu_threads)
        if cpu is None:       # reference harness not built: the port is the only baseline
            cpu, cpu_port = cpu_port, None

    out = {
        "metric": "Mcell-updates/s", "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded conduction profile + 1e-3 noise)",
        "config": {"workload": f"{args.config}: RB cell {p['nxm']}x{p['nym']}x{p['nzm']} istr3={p['istr3']} "
                               f"str3={p['str3']} Ra={p['ray']:g} Pr={p['pra']:g} RK3 CFL dt", "cells": cells},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_port": cpu_port,
        "substep": {"ms": round(substep_ms, 3), "alg_GBps_65words": round(substep_alg, 1),
                    "frac_hbm_65words": round(substep_alg / HBM_PEAK_GBS, 4)},
        "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in kern.items()},
        "max_divergence": divmax,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Writes the golden fixtures of tests/golden/ from the reference routines.

Runs oracle/_ref/ref_harness (reference AFiD Fortran compiled unchanged from
the reference tree by oracle/Makefile, with the oracle's Poisson solve in
place of the FFTW-dependent SolvePressureCorrection) on small seeded inputs
and stores inputs and outputs as compressed .npz data:

  grid_<case>.npz    CreateGrid outputs (mode 1)
  stages_<case>.npz  per-routine stages A-E from random fields (mode 2)
  march_<case>.npz   10 TimeMarcher steps from the seeded linear+noise IC (mode 3)

Usage: python tests/golden/make_golden.py  (needs the reference tree and `make -C oracle ref`).
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle.oracle import Oracle, ref_available, run_ref  # noqa: E402

GRID_CASES = {
    "tanh": dict(nxm=16, nym=12, nzm=10, nsst=3, alx3=1.0, istr3=4, str3=4.0, ylen=2.0, zlen=2.0, ray=1e6, pra=1.0,
                 inslws=1, inslwn=1),
    "cheb": dict(nxm=24, nym=12, nzm=10, nsst=3, alx3=1.0, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e8, pra=1.0,
                 inslws=1, inslwn=1),
    "unif_slip": dict(nxm=16, nym=12, nzm=10, nsst=3, alx3=1.0, istr3=0, str3=0.0, ylen=1.0, zlen=1.0, ray=1e5,
                      pra=0.7, inslws=0, inslwn=1),
}
STAGE_CASES = {
    "rk_sub1": (dict(GRID_CASES["tanh"], ylen=2.0, zlen=1.5, ray=1e5, pra=0.7, str3=3.0), 1, 2e-3),
    "rk_sub2": (dict(GRID_CASES["tanh"], ylen=2.0, zlen=1.5, ray=1e5, pra=0.7, str3=3.0), 2, 2e-3),
    "ab2": (dict(GRID_CASES["cheb"], nsst=1), 1, 1e-3),
}
MARCH_CASES = {
    "fixed_dt": (dict(nxm=16, nym=16, nzm=12, nsst=3, alx3=1.0, istr3=4, str3=4.0, ylen=2.0, zlen=2.0, ray=1e6,
                      pra=1.0, inslws=1, inslwn=1), 0, 2e-3),
    "cfl_dt": (dict(nxm=16, nym=16, nzm=12, nsst=3, alx3=1.0, istr3=6, str3=6.0, ylen=1.0, zlen=1.0, ray=1e7,
                    pra=0.7, inslws=1, inslwn=1), 1, 1e-4),
}


def halo_shape(p):
    return (p["nzm"] + 2, p["nym"] + 2, p["nxm"] + 1)


def plain_shape(p, n1=None):
    return (p["nzm"], p["nym"], n1 if n1 else p["nxm"] + 1)


def make_grid(name, p, wd):
    run_ref(wd, p, mode=1)
    g = np.fromfile(os.path.join(wd, "grid_out.bin"))
    np.savez_compressed(os.path.join(HERE, f"grid_{name}.npz"),
                        raw=g, **{k: np.array(v) for k, v in p.items()})


def make_stages(name, p, ns, dt, wd, seed=5):
    rng = np.random.default_rng(seed)
    hs, ps = halo_shape(p), plain_shape(p)
    F = {n: rng.standard_normal(hs) for n in ("vx", "vy", "vz", "temp", "pr")}
    F.update({n: rng.standard_normal(ps) for n in ("rux", "ruy", "ruz", "rutemp", "rhs")})
    dphin = rng.standard_normal(plain_shape(p, p["nxm"]))
    with open(os.path.join(wd, "fields_in.bin"), "wb") as f:
        for n in ("vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp", "rhs"):
            f.write(F[n].tobytes())
    dphin.tofile(os.path.join(wd, "dph_in.bin"))
    run_ref(wd, p, mode=2, nsteps=ns, dt=dt)
    out = {}
    nh, npl, nm = int(np.prod(hs)), int(np.prod(ps)), int(np.prod(plain_shape(p, p["nxm"])))
    A = np.fromfile(os.path.join(wd, "stageA.bin"))
    out["qcap"], out["dph_expl"] = A[:npl].reshape(ps), A[npl:npl + nm].reshape(plain_shape(p, p["nxm"]))
    out["dq"], out["hro"] = A[npl + nm:2 * npl + nm].reshape(ps), A[2 * npl + nm:].reshape(ps)
    B = np.fromfile(os.path.join(wd, "stageB.bin"))
    off = 0
    for n in ("vx", "vy", "vz", "temp"):
        out[n + "_impl"] = B[off:off + nh].reshape(hs)
        off += nh
    for n in ("rhs", "rux", "ruy", "ruz", "rutemp"):
        out[n + "_impl"] = B[off:off + npl].reshape(ps)
        off += npl
    out["dph_div"] = np.fromfile(os.path.join(wd, "stageC.bin")).reshape(plain_shape(p, p["nxm"]))
    D = np.fromfile(os.path.join(wd, "stageD.bin"))
    for q, n in enumerate(("vx", "vy", "vz", "pr")):
        out[n + "_corr"] = D[q * nh:(q + 1) * nh].reshape(hs)
    E = np.fromfile(os.path.join(wd, "stageE.bin"))
    out["diag"] = E
    out["nu_vol"] = float(open(os.path.join(wd, "nu_vol.out")).read().split()[1])
    nud = open(os.path.join(wd, "nu_diss.out")).read().split()
    out["nu_diss"] = np.array([float(nud[1]), float(nud[2])])
    nup = open(os.path.join(wd, "nu_plate.out")).read().split()   # format 4(1x,e14.6): 6 digits
    out["nu_plate"] = np.array([float(nup[1]), float(nup[2])])
    for fn in ("nu_vol.out", "nu_diss.out", "nu_plate.out", "rms_vel.out"):
        os.remove(os.path.join(wd, fn))
    np.savez_compressed(os.path.join(HERE, f"stages_{name}.npz"), ns=ns, dt=dt, dph_in=dphin,
                        **{"in_" + k: v for k, v in F.items()}, **out, **{k: np.array(v) for k, v in p.items()})


def make_march(name, p, idtv, dt, wd, seed=7, nsteps=10):
    o = Oracle(**p)
    o.init_linear_noise(seed, 1e-3)
    with open(os.path.join(wd, "fields_in.bin"), "wb") as f:
        for n in ("vx", "vy", "vz", "temp", "pr", "rux", "ruy", "ruz", "rutemp", "rhs"):
            f.write(o.arr(n).tobytes())
    run_ref(wd, p, mode=3, nsteps=nsteps, dt=dt, idtv=idtv, dtmin=1e-4, dtmax=5e-3, limitcfl=1.2)
    diag = np.fromfile(os.path.join(wd, "march_diag.bin")).reshape(-1, 3)
    M = np.fromfile(os.path.join(wd, "march_out.bin"))
    hs = halo_shape(p)
    nh = int(np.prod(hs))
    out = {n: M[q * nh:(q + 1) * nh].reshape(hs) for q, n in enumerate(("vx", "vy", "vz", "temp", "pr"))}
    np.savez_compressed(os.path.join(HERE, f"march_{name}.npz"), seed=seed, idtv=idtv, dt=dt, nsteps=nsteps,
                        dtmin=1e-4, dtmax=5e-3, limitcfl=1.2, diag=diag, **out,
                        **{k: np.array(v) for k, v in p.items()})


def main():
    if not ref_available():
        sys.exit("reference harness not built: make -C oracle ref")
    with tempfile.TemporaryDirectory() as wd:
        for n, p in GRID_CASES.items():
            make_grid(n, p, wd)
        for n, (p, ns, dt) in STAGE_CASES.items():
            make_stages(n, p, ns, dt, wd)
        for n, (p, idtv, dt) in MARCH_CASES.items():
            make_march(n, p, idtv, dt, wd)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()

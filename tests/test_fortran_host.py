"""The Fortran host (afid_amd/lib/afid_run): bou.in reader, bind(C) layout, and
on the GPU a full main-loop run compared with the oracle's TimeMarcher.

The host follows main.F90's loop (ntst+1 steps, CFL-limited dt with dtmin for
the first two steps) and ends with errorcode 555; the final fields it writes
must match the oracle stepped with the same dt logic to rel-L2 <= 1e-11
(pr modulo its mean), through both the fused substep and the per-routine
entry points.  With STATON and SLABDUMP set the host also writes
stafield_master.h5 (one sample per output time after tsta), stst3locs.out and
the stst3/ slab files at each output time.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import demean, h5_values, rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "afid_amd", "lib", "afid_run")

BOU = """ NXM  NYM  NZM   NSST(>1=RK,else AB) NREAD
 {nxm} {nym} {nzm}       3                   0
 NTST   WALLTIMEMAX    TOUT      TMAX     IRESET
 {ntst}     83000          1e-3      50.0      0
 ALX3D      ISTR3   STR3
 1.0000       {istr3}     {str3}
 YLEN ZLEN
 {ylen}   {zlen}
   RA   PRA      DT      RESID    CFLMAX
 {ray}  {pra}   {dt}   1.e-2    1.2
 STATON  BALANCEON  TSTA  STAREAD
 1        1            0    0
 inslws   inslwn
  1          1
 IDTV(0=dt fixed,1=var.)   DTMIN(for dt var.)  DTMAX(for dt var.)     VLIM
   {idtv}                       {dtmin}                {dtmax}                    1e3
 SLABDUMP(STST3)
   {stst3}
"""

CASE = dict(nxm=24, nym=16, nzm=20, istr3=4, str3=3.0, ylen=2.0, zlen=1.5, ray=1e6, pra=1.0)


def write_bou(path, ntst=3, idtv=1, dt=2e-3, dtmin=1e-4, dtmax=5e-3, stst3=0, **c):
    q = dict(CASE)
    q.update(c)
    with open(path, "w") as f:
        f.write(BOU.format(ntst=ntst, idtv=idtv, dt=dt, dtmin=dtmin, dtmax=dtmax, stst3=stst3, **q))


def need_run():
    if not os.path.exists(RUN):
        pytest.skip("afid_run not built (python -c 'import __graft_entry__ as g; g.build()')")


def test_bind_c_params_layout():
    need_run()
    from afid_amd import engine
    out = subprocess.run([RUN, "--abi-check"], capture_output=True, text=True, check=True).stdout
    assert f"sizeof(afid_params)={ctypes.sizeof(engine._Params)}" in out


def test_bou_in_reader(tmp_path):
    need_run()
    write_bou(tmp_path / "bou.in", ntst=7, nxm=32, nym=48, nzm=40, istr3=6, str3=6.0)
    out = subprocess.run([RUN, "bou.in", "--parse-only"], cwd=tmp_path, capture_output=True, text=True,
                         check=True).stdout
    lines = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    assert lines["grid"] == ["32", "48", "40", "3", "0"]
    assert lines["ntst"][0] == "7"
    assert lines["stretch"][1] == "6" and float(lines["stretch"][2]) == 6.0
    assert [float(v) for v in lines["physics"][:2]] == [1e6, 1.0]
    assert lines["walls"] == ["1", "1"]
    assert float(lines["timestep"][1]) == 1e-4 and float(lines["timestep"][2]) == 5e-3


@pytest.mark.gpu
@pytest.mark.parametrize("routines", [False, True], ids=["fused", "routines"])
def test_fortran_main_loop_matches_oracle(tmp_path, routines):
    need_run()
    from oracle.oracle import Oracle
    ntst, seed = 3, 4
    write_bou(tmp_path / "bou.in", ntst=ntst)
    cmd = [RUN, "bou.in", "--seed", str(seed), "--out", "."] + (["--routines"] if routines else [])
    print("afid_run", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errorcode 555" in r.stdout
    o = Oracle(nsst=3, alx3=1.0, **CASE)
    o.init_linear_noise(seed, 1e-3)
    for n in range(ntst + 1):
        o.step(n, 1, 2e-3, 1e-4, 5e-3, 1.2)
    nx, nym, nzm = CASE["nxm"] + 1, CASE["nym"], CASE["nzm"]
    got = {}
    for nm in ("vx", "vy", "vz", "temp", "pr"):
        a = np.fromfile(tmp_path / f"{nm}.bin", dtype=np.float64)
        got[nm] = a.reshape(nzm, nym, nx)           # Fortran (nx, nym, nzm) read as C (nzm, nym, nx)
    for nm in ("vx", "vy", "vz", "temp"):
        assert rel_l2(got[nm], o.interior(nm)) < 1e-11, nm
    nxm = CASE["nxm"]
    assert rel_l2(demean(got["pr"][:, :, :nxm]), demean(o.interior("pr")[:, :, :nxm])) < 1e-11
    t = float(r.stdout.split("errorcode 555 time")[1].split()[0])
    assert t == pytest.approx(o.scalar("time"), rel=1e-12)
    assert os.path.exists(tmp_path / "nu_plate.out") and os.path.exists(tmp_path / "nu_vol.out")


@pytest.mark.gpu
def test_fortran_stats_and_slabs(tmp_path):
    need_run()
    from afid_amd import engine
    from oracle.oracle import Oracle
    write_bou(tmp_path / "bou.in", ntst=4, stst3=1)
    xslab = [0.002, 0.004, 0.008, 0.02, 0.04, 0.08, 0.2, 0.5]
    with open(tmp_path / "stst3.in", "w") as f:
        f.write(" PROBEZ1\n " + "  ".join(str(v) for v in xslab) + "\n")
    print("afid_run stats+slabs", flush=True)
    r = subprocess.run([RUN, "bou.in", "--seed", "2"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "errorcode 555" in r.stdout
    o = Oracle(nsst=3, alx3=1.0, **CASE)
    xm = o.arr("xm")[1:CASE["nxm"] + 1]
    ks = [int(v) for v in open(tmp_path / "stst3locs.out").read().split()]
    assert ks == list(engine.slab_levels(xm, np.array(xslab)))
    nout = len(open(tmp_path / "nu_plate.out").read().splitlines())
    assert nout >= 2
    assert h5_values(tmp_path / "stafield_master.h5", "averaging_time")[0] == nout
    assert len(os.listdir(tmp_path / "stst3")) % 36 == 0 and len(os.listdir(tmp_path / "stst3")) >= 36
    np.testing.assert_array_equal(h5_values(tmp_path / "stafield_master.h5", "X_cordin"), xm)


@pytest.mark.gpu
def test_fortran_mixed_precision_run(tmp_path):
    """afid_run --mixed: the mixed-precision mode from the Fortran host; its final
    fields stay within the stated mixed-vs-FP64 error of the FP64 run."""
    need_run()
    out = {}
    for mode in ("fp64", "mixed"):
        d = tmp_path / mode
        d.mkdir()
        write_bou(d / "bou.in", ntst=3)
        cmd = [RUN, "bou.in", "--seed", "4", "--out", "."] + (["--mixed"] if mode == "mixed" else [])
        print("afid_run", " ".join(cmd[1:]), flush=True)
        r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "errorcode 555" in r.stdout
        nx, nym, nzm = CASE["nxm"] + 1, CASE["nym"], CASE["nzm"]
        out[mode] = {nm: np.fromfile(d / f"{nm}.bin", dtype=np.float64).reshape(nzm, nym, nx)
                     for nm in ("vx", "vy", "vz", "temp")}
    for nm in ("vx", "vy", "vz"):
        assert rel_l2(out["mixed"][nm], out["fp64"][nm]) < 2e-5, nm
    assert rel_l2(out["mixed"]["temp"], out["fp64"]["temp"]) < 1e-9

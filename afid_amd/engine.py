"""Host-side mirror of AFiD's substep interface over the C ABI (include/afid.h).

The reference exposes argument-less Fortran subroutines acting on module
globals (param.F90 / local_arrays).  Here a ``Cell`` owns one device context
(the role of those modules) and offers the same routine names:

    ExplicitTermsVX/VY/VZ/Temp, ImplicitAndUpdateVX/VY/VZ/Temp,
    CalcLocalDivergence, SolvePressureCorrection, CorrectVelocity,
    CorrectPressure, update_halo, TimeMarcher,
    CalcMaxCFL, CheckDivergence, CalcPlateNu, GlobalQuantities, CalcDissipationNu

Arrays cross the boundary in the reference's Fortran layout (x fastest);
numpy views here are indexed [i, j, k] (z, y, x).  The library is loaded from
afid_amd/lib/libafid_hip.so; there is no fallback: if it is missing or a HIP
call fails, an AfidError is raised.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libafid_hip.so")
IO_LIB_PATH = os.path.join(HERE, "lib", "libafid_io.so")

AFID_VX, AFID_VY, AFID_VZ, AFID_TEMP, AFID_PR = range(5)
AFID_RUX, AFID_RUY, AFID_RUZ, AFID_RUTEMP, AFID_DPH, AFID_QCAP, AFID_DQ, AFID_HRO, AFID_RHS = range(5, 14)
FIELD_IDS = {"vx": AFID_VX, "vy": AFID_VY, "vz": AFID_VZ, "temp": AFID_TEMP, "pr": AFID_PR,
             "rux": AFID_RUX, "ruy": AFID_RUY, "ruz": AFID_RUZ, "rutemp": AFID_RUTEMP, "dph": AFID_DPH,
             "qcap": AFID_QCAP, "dq": AFID_DQ, "hro": AFID_HRO, "rhs": AFID_RHS}


class AfidError(RuntimeError):
    """A nonzero return code of the library; .code holds it (0 when none applies)."""

    def __init__(self, msg: str, code: int = 0):
        super().__init__(msg)
        self.code = code


class _Params(ctypes.Structure):
    _fields_ = [("nxm", ctypes.c_int), ("nym", ctypes.c_int), ("nzm", ctypes.c_int),
                ("nsst", ctypes.c_int), ("istr3", ctypes.c_int), ("str3", ctypes.c_double),
                ("alx3", ctypes.c_double), ("ylen", ctypes.c_double), ("zlen", ctypes.c_double),
                ("ray", ctypes.c_double), ("pra", ctypes.c_double), ("inslws", ctypes.c_int),
                ("inslwn", ctypes.c_int), ("tempbp", ctypes.c_double), ("temptp", ctypes.c_double),
                ("precision", ctypes.c_int)]

PREC_FP64, PREC_MIXED = 0, 1

# afid_metrics members in declaration order: 1-D arrays with their lengths
# ("nx", "nxm", "nzm" or "nym"); an omitted one keeps the library's own grid.
METRIC_NAMES = (("xc", "nx"), ("xm", "nx"), ("g3rc", "nx"), ("g3rm", "nx"), ("udx3c", "nx"), ("udx3m", "nx"),
                ("ap3ck", "nx"), ("ac3ck", "nx"), ("am3ck", "nx"), ("ap3sk", "nx"), ("ac3sk", "nx"),
                ("am3sk", "nx"), ("ap3ssk", "nx"), ("ac3ssk", "nx"), ("am3ssk", "nx"),
                ("amphk", "nxm"), ("acphk", "nxm"), ("apphk", "nxm"), ("ak1", "nzm"), ("ak2", "nym"))


class _Metrics(ctypes.Structure):
    _fields_ = [(nm, ctypes.POINTER(ctypes.c_double)) for nm, _ in METRIC_NAMES]


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libafid_hip.so; raises AfidError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AfidError(f"{path} not found: build it with afid_amd.build.build() (hipcc, gfx950)")
    L = ctypes.CDLL(path)
    vp, i, d, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
    pd = ctypes.POINTER(d)
    sig = {
        "afid_create": ([ctypes.POINTER(vp), ctypes.POINTER(_Params), ctypes.POINTER(_Metrics)], i),
        "afid_destroy": ([vp], i),
        "afid_strerror": ([i], ctypes.c_char_p),
        "afid_last_hip_error": ([vp], i),
        "afid_get_dims": ([vp, ctypes.POINTER(i64)], i),
        "afid_get_metric": ([vp, ctypes.c_char_p, pd, i], i),
        "afid_upload": ([vp, i, pd, i, i, i, i], i),
        "afid_download": ([vp, i, pd, i, i, i, i], i),
        "afid_init_linear_noise": ([vp, ctypes.c_uint64, d], i),
        "afid_set_coefficients": ([vp, d, d, d, d, d], i),
        "afid_set_substep": ([vp, i, d], i),
        "afid_substep": ([vp, i, d], i),
        "afid_time_marcher": ([vp, d], i),
        "afid_step": ([vp, i, i, d, d, d, d, pd, ctypes.POINTER(i)], i),
        "afid_steps": ([vp, i, i, i, d, d, d, d, pd, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "afid_steps_async": ([vp, i, i, i, d, d, d, d], i),
        "afid_steps_more": ([vp, i], i),
        "afid_steps_query": ([vp, ctypes.POINTER(i)], i),
        "afid_steps_wait": ([vp, pd, ctypes.POINTER(i), ctypes.POINTER(i)], i),
        "afid_steps_cfl": ([vp, pd], i),
        "afid_batch_steps": ([ctypes.POINTER(vp), i, i], i),
        "afid_estimate_bytes": ([ctypes.c_void_p, ctypes.POINTER(i64)], i),
        "afid_tune": ([vp, ctypes.POINTER(i)], i),
        "afid_get_time": ([vp, pd], i),
        "afid_set_time": ([vp, d], i),
        "afid_get_stale": ([vp, pd], i),
        "afid_calc_stats": ([vp], i),
        "afid_get_stats": ([vp, pd, ctypes.POINTER(i)], i),
        "afid_reset_stats": ([vp], i),
        "afid_calc_stats_noalloc": ([vp, vp], i),
        "afid_get_slabs": ([vp, ctypes.POINTER(i), i, pd], i),
        "afid_interp_field": ([vp, i, pd, i, i, i, i, d], i),
        "afid_set_stale": ([vp, pd], i),
        "afid_calc_max_cfl": ([vp, pd], i),
        "afid_check_divergence": ([vp, pd], i),
        "afid_calc_plate_nu": ([vp, pd, pd], i),
        "afid_global_quantities": ([vp, pd], i),
        "afid_calc_dissipation_nu": ([vp, pd, pd, pd, pd], i),
        "afid_update_halo": ([vp, i], i),
        "afid_pressure_spectrum": ([vp, i, pd], i),
        "afid_profile_enable": ([vp, i], i),
        "afid_profile_read": ([vp, ctypes.c_char_p, i, pd, ctypes.POINTER(i64), i], i),
        "afid_profile_reset": ([vp], i),
        "afid_synchronize": ([vp], i),
    }
    for name in ("explicit_terms_vx", "explicit_terms_vy", "explicit_terms_vz", "explicit_terms_temp",
                 "implicit_and_update_vx", "implicit_and_update_vy", "implicit_and_update_vz",
                 "implicit_and_update_temp", "calc_local_divergence", "solve_pressure_correction",
                 "correct_velocity", "correct_pressure"):
        sig["afid_" + name] = ([vp], i)
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


_io = None


def load_io_library(path: str = IO_LIB_PATH) -> ctypes.CDLL:
    """Load libafid_io.so (continuation files); raises AfidError if not built."""
    global _io
    if _io is not None:
        return _io
    load_library()
    if not os.path.exists(path):
        raise AfidError(f"{path} not found: build it with afid_amd.build.build()")
    L = ctypes.CDLL(path)
    vp, i, pd, cp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_char_p
    pi = ctypes.POINTER(ctypes.c_int)
    sig = {"afid_write_continuation": ([vp, cp], i), "afid_read_continuation": ([vp, cp], i),
           "afid_write_stats": ([vp, cp, i, i], i), "afid_slab_levels": ([pd, i, pd, pi], i),
           "afid_write_slabs": ([vp, cp, pi, ctypes.c_double, ctypes.c_double], i),
           "afid_h5_write_field": ([cp, cp, pd, i, i, i], i), "afid_h5_read_field": ([cp, cp, pd, i, i, i], i)}
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _io = L
    return L


def exported_symbols(header: str = "afid.h") -> list[str]:
    """Entry points declared in include/<header> (for the export check)."""
    hdr = os.path.join(os.path.dirname(HERE), "include", header)
    import re
    names = re.findall(r"^(?:int|const char \*)\s*(afid_\w+)\s*\(", open(hdr).read(), flags=re.M)
    return names


STAT_NAMES = ("vx_mean", "vy_mean", "vz_mean", "temp_mean", "vx_rms", "vy_rms", "vz_rms", "temp_rms",
              "tempvx_mean", "disste", "dissth")


def slab_levels(xm: np.ndarray, xslab: np.ndarray) -> np.ndarray:
    """InitializeSlabDump: kslab(1..9) for xm(1..nxm) and xslab(2..9)."""
    xm = np.ascontiguousarray(xm, dtype=np.float64)
    xs = np.ascontiguousarray(xslab, dtype=np.float64)
    out = (ctypes.c_int * 9)()
    rc = load_io_library().afid_slab_levels(_dp(xm), len(xm), _dp(xs), out)
    if rc:
        raise AfidError(f"slab_levels: {rc}", rc)
    return np.array(out[:], dtype=np.int32)


def h5_write_field(path: str, dset: str, a: np.ndarray) -> None:
    """One FP64 dataset of Fortran shape (n1, n2, n3) from a numpy [i, j, k] array."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n3, n2, n1 = a.shape
    rc = load_io_library().afid_h5_write_field(path.encode(), dset.encode(), _dp(a), n1, n2, n3)
    if rc:
        raise AfidError(f"h5_write_field: {load_library().afid_strerror(rc).decode()} ({rc})", rc)


def h5_read_field(path: str, dset: str, shape: tuple) -> np.ndarray:
    """Reads a dataset of Fortran shape (n1, n2, n3) into a numpy [i, j, k] array of `shape`."""
    a = np.zeros(shape)
    n3, n2, n1 = shape
    rc = load_io_library().afid_h5_read_field(path.encode(), dset.encode(), _dp(a), n1, n2, n3)
    if rc:
        raise AfidError(f"h5_read_field: {load_library().afid_strerror(rc).decode()} ({rc})", rc)
    return a


@dataclass
class CellParams:
    """bou.in subset (ReadInputFile.F90:16-36)."""
    nxm: int
    nym: int
    nzm: int
    nsst: int = 3
    istr3: int = 0
    str3: float = 0.0
    alx3: float = 1.0
    ylen: float = 1.0
    zlen: float = 1.0
    ray: float = 1e6
    pra: float = 1.0
    inslws: int = 1
    inslwn: int = 1
    tempbp: float = 1.0
    temptp: float = 0.0
    precision: int = PREC_FP64     # PREC_MIXED: FP32 histories and FFT spectrum

    def to_c(self) -> _Params:
        return _Params(self.nxm, self.nym, self.nzm, self.nsst, self.istr3, self.str3, self.alx3,
                       self.ylen, self.zlen, self.ray, self.pra, self.inslws, self.inslwn,
                       self.tempbp, self.temptp, self.precision)

    @property
    def cells(self) -> int:
        return self.nxm * self.nym * self.nzm


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def estimate_bytes(params: "CellParams") -> int:
    """Device bytes a cell of these parameters holds at most."""
    L = load_library()
    cp = params.to_c()
    out = ctypes.c_int64()
    rc = L.afid_estimate_bytes(ctypes.byref(cp), ctypes.byref(out))
    if rc:
        raise AfidError(f"estimate_bytes: {L.afid_strerror(rc).decode()} ({rc})", rc)
    return out.value


def batch_steps(cells, nsteps: int):
    """Queues nsteps more steps on every cell, interleaved (each had steps_async)."""
    L = load_library()
    hs = (ctypes.c_void_p * len(cells))(*[c.h.value for c in cells])
    rc = L.afid_batch_steps(hs, len(cells), nsteps)
    if rc:
        raise AfidError(f"batch_steps: {L.afid_strerror(rc).decode()} ({rc})", rc)


class Cell:
    """One Rayleigh-Benard cell resident on the GPU."""

    def __init__(self, params: CellParams, metrics: Optional[dict] = None):
        """metrics: caller-supplied 1-D grid arrays by param.F90 name (1-based
        arrays given from index 1, as afid_metrics takes them); they replace the
        library's own CreateGrid / InitPressureSolver values."""
        self.L = load_library()
        self.params = params
        self.h = ctypes.c_void_p()
        cp = params.to_c()
        mp = None
        if metrics:
            sizes = {"nx": params.nxm + 1, "nxm": params.nxm, "nym": params.nym, "nzm": params.nzm}
            unknown = set(metrics) - {nm for nm, _ in METRIC_NAMES}
            if unknown:
                raise ValueError(f"unknown metric arrays: {sorted(unknown)}")
            self._metric_arrays = {}
            m = _Metrics()
            for nm, n in METRIC_NAMES:
                if nm in metrics:
                    a = np.ascontiguousarray(metrics[nm], dtype=np.float64)
                    if a.shape != (sizes[n],):
                        raise ValueError(f"metric {nm}: {sizes[n]} values expected, got {a.shape}")
                    self._metric_arrays[nm] = a
                    setattr(m, nm, _dp(a))
            mp = ctypes.byref(m)
        self._check(self.L.afid_create(ctypes.byref(self.h), ctypes.byref(cp), mp), "afid_create")
        dims = (ctypes.c_int64 * 6)()
        self._check(self.L.afid_get_dims(self.h, dims), "afid_get_dims")
        self.nx, self.nxm, self.nym, self.nzm, self.ldx, self.device_bytes = list(dims)

    def held_bytes(self) -> int:
        """Device bytes the context holds now (buffers allocated so far)."""
        dims = (ctypes.c_int64 * 6)()
        self._check(self.L.afid_get_dims(self.h, dims), "afid_get_dims")
        return int(dims[5])

    # ---- plumbing ---------------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.L.afid_strerror(rc).decode()
            hip = self.L.afid_last_hip_error(self.h) if self.h else 0
            raise AfidError(f"{what}: {msg} (code {rc}, hip {hip})", rc)

    def close(self):
        if self.h:
            self.L.afid_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def metric(self, name: str, n: int) -> np.ndarray:
        out = np.zeros(n)
        self._check(self.L.afid_get_metric(self.h, name.encode(), _dp(out), n), f"metric {name}")
        return out

    # ---- data exchange ----------------------------------------------------------
    def upload(self, name: str, a: np.ndarray, halo: bool = False):
        """a: (nzm[+2], nym[+2], n1) C-order = Fortran (n1, nym, nzm) with x fastest."""
        fid = FIELD_IDS[name]
        a = np.ascontiguousarray(a, dtype=np.float64)
        j0 = i0 = 0 if halo else 1
        self._check(self.L.afid_upload(self.h, fid, _dp(a), a.shape[2], a.shape[1], j0, i0), f"upload {name}")

    def download(self, name: str, halo: bool = False) -> np.ndarray:
        fid = FIELD_IDS[name]
        n1 = self.nxm if name == "dph" else self.nx
        ny, nz = (self.nym + 2, self.nzm + 2) if halo else (self.nym, self.nzm)
        a = np.zeros((nz, ny, n1))
        j0 = i0 = 0 if halo else 1
        self._check(self.L.afid_download(self.h, fid, _dp(a), n1, ny, j0, i0), f"download {name}")
        return a

    def init_linear_noise(self, seed: int, amp: float = 1e-3):
        self._check(self.L.afid_init_linear_noise(self.h, seed, amp), "init_linear_noise")

    def synchronize(self):
        self._check(self.L.afid_synchronize(self.h), "synchronize")

    # ---- reference routine names -------------------------------------------------
    def set_coefficients(self, al, ga, ro, dt, beta):
        self._check(self.L.afid_set_coefficients(self.h, al, ga, ro, dt, beta), "set_coefficients")

    def set_substep(self, ns: int, dt: float):
        self._check(self.L.afid_set_substep(self.h, ns, dt), "set_substep")

    def _call(self, name):
        self._check(getattr(self.L, "afid_" + name)(self.h), name)

    def ExplicitTermsVX(self): self._call("explicit_terms_vx")
    def ExplicitTermsVY(self): self._call("explicit_terms_vy")
    def ExplicitTermsVZ(self): self._call("explicit_terms_vz")
    def ExplicitTermsTemp(self): self._call("explicit_terms_temp")
    def ImplicitAndUpdateVX(self): self._call("implicit_and_update_vx")
    def ImplicitAndUpdateVY(self): self._call("implicit_and_update_vy")
    def ImplicitAndUpdateVZ(self): self._call("implicit_and_update_vz")
    def ImplicitAndUpdateTemp(self): self._call("implicit_and_update_temp")
    def CalcLocalDivergence(self): self._call("calc_local_divergence")
    def SolvePressureCorrection(self): self._call("solve_pressure_correction")
    def CorrectVelocity(self): self._call("correct_velocity")
    def CorrectPressure(self): self._call("correct_pressure")

    def pressure_spectrum(self, stage: int) -> np.ndarray:
        """Spectrum after `stage` steps of SolvePressureCorrection, indexed [n, m, k]."""
        nymh = self.nym // 2 + 1
        out = np.zeros((self.nzm, nymh, self.nxm), dtype=np.complex128)
        self._check(self.L.afid_pressure_spectrum(self.h, stage, _dp(out.view(np.float64))), "pressure_spectrum")
        return out

    def update_halo(self, name: str):
        self._check(self.L.afid_update_halo(self.h, FIELD_IDS[name]), "update_halo")

    def substep(self, ns: int, dt: float):
        self._check(self.L.afid_substep(self.h, ns, dt), "substep")

    def TimeMarcher(self, dt: float):
        self._check(self.L.afid_time_marcher(self.h, dt), "TimeMarcher")

    def step(self, ntime: int, idtv: int, dt: float, dtmin: float = 1e-8, dtmax: float = 1e-2,
             limitcfl: float = 1.2):
        used = ctypes.c_double()
        err = ctypes.c_int()
        self._check(self.L.afid_step(self.h, ntime, idtv, dt, dtmin, dtmax, limitcfl, ctypes.byref(used),
                                     ctypes.byref(err)), "step")
        return err.value, used.value

    def steps(self, ntime0: int, nsteps: int, idtv: int, dt: float, dtmin: float = 1e-8, dtmax: float = 1e-2,
              limitcfl: float = 1.2):
        """nsteps main.F90 steps on the device (one captured graph per step);
        returns (errorcode, dt of the last step run, steps run)."""
        used = ctypes.c_double()
        err, done = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.afid_steps(self.h, ntime0, nsteps, idtv, dt, dtmin, dtmax, limitcfl, ctypes.byref(used),
                                      ctypes.byref(err), ctypes.byref(done)), "steps")
        return err.value, used.value, done.value

    def steps_async(self, ntime0: int, nsteps: int, idtv: int, dt: float, dtmin: float = 1e-8,
                    dtmax: float = 1e-2, limitcfl: float = 1.2):
        """Queues nsteps device steps on this cell's stream and returns at once."""
        self._check(self.L.afid_steps_async(self.h, ntime0, nsteps, idtv, dt, dtmin, dtmax, limitcfl),
                    "steps_async")

    def steps_cfl(self) -> float:
        """CalcMaxCFL of the fields after the last device step, as the device step
        control computed it (folded into the last correction in FP64)."""
        v = ctypes.c_double()
        self._check(self.L.afid_steps_cfl(self.h, ctypes.byref(v)), "steps_cfl")
        return v.value

    def tune(self) -> int:
        """Autotunes the stencil pass's z planes per block on this cell (call before
        setting the initial state); returns the chosen number of planes."""
        zm = ctypes.c_int()
        self._check(self.L.afid_tune(self.h, ctypes.byref(zm)), "tune")
        return zm.value

    def steps_more(self, nsteps: int):
        self._check(self.L.afid_steps_more(self.h, nsteps), "steps_more")

    def busy(self) -> bool:
        b = ctypes.c_int()
        self._check(self.L.afid_steps_query(self.h, ctypes.byref(b)), "steps_query")
        return bool(b.value)

    def steps_wait(self):
        """(errorcode, dt of the last step, steps run since steps_async)."""
        used = ctypes.c_double()
        err, done = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.afid_steps_wait(self.h, ctypes.byref(used), ctypes.byref(err), ctypes.byref(done)),
                    "steps_wait")
        return err.value, used.value, done.value

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        self._check(self.L.afid_get_time(self.h, ctypes.byref(t)), "get_time")
        return t.value

    @time.setter
    def time(self, t: float):
        self._check(self.L.afid_set_time(self.h, t), "set_time")

    # ---- statistics and slabs (StatRoutines.F90 / SlabDumpRoutines.F90) ---------
    def CalcStats(self):
        self._check(self.L.afid_calc_stats(self.h), "CalcStats")

    def stats(self):
        """Accumulated statistics {name: array(nxm)} and the number of samples."""
        out = np.zeros(11 * self.nxm)
        n = ctypes.c_int()
        self._check(self.L.afid_get_stats(self.h, _dp(out), ctypes.byref(n)), "get_stats")
        return {nm: out[q * self.nxm:(q + 1) * self.nxm] for q, nm in enumerate(STAT_NAMES)}, n.value

    def reset_stats(self):
        self._check(self.L.afid_reset_stats(self.h), "reset_stats")

    def get_stale(self):
        """The stale rhs(1) value of every column (the vz increment at k = 1 of
        the last substep), as (nzm, nym)."""
        out = np.zeros(self.nym * self.nzm)
        self._check(self.L.afid_get_stale(self.h, _dp(out)), "get_stale")
        return out.reshape(self.nzm, self.nym)

    def CalcStatsNoAlloc(self, acc_ptr: int):
        """afid_calc_stats_noalloc: adds this sample's nine plane-average
        profiles into the caller's device array at acc_ptr (9*nxm doubles);
        allocates nothing and returns without synchronising."""
        self._check(self.L.afid_calc_stats_noalloc(self.h, ctypes.c_void_p(acc_ptr)), "CalcStatsNoAlloc")

    def WriteStats(self, directory: str = ".", readstats: bool = False, disscal: bool = True):
        self._check(load_io_library().afid_write_stats(self.h, directory.encode(), int(readstats), int(disscal)),
                    "WriteStats")

    def slabs(self, kslab) -> np.ndarray:
        """Planes [m, q, i, j] at levels kslab (1-based): q = (vx(k)+vx(k+1))/2, vy, vz, temp."""
        ks = (ctypes.c_int * len(kslab))(*[int(k) for k in kslab])
        out = np.zeros((len(kslab), 4, self.nzm, self.nym))
        self._check(self.L.afid_get_slabs(self.h, ks, len(kslab), _dp(out)), "get_slabs")
        return out

    def SlabDumper(self, directory: str, kslab, time: float, tout: float):
        ks = (ctypes.c_int * 9)(*[int(k) for k in kslab])
        self._check(load_io_library().afid_write_slabs(self.h, directory.encode(), ks, time, tout), "SlabDumper")

    # ---- continuation files (WriteFlowField.F90 / ReadFlowField.F90) -------------
    def WriteFlowField(self, directory: str = "."):
        self._check(load_io_library().afid_write_continuation(self.h, directory.encode()), "WriteFlowField")

    def ReadFlowField(self, directory: str = "."):
        self._check(load_io_library().afid_read_continuation(self.h, directory.encode()), "ReadFlowField")

    def interp(self, name: str, old: np.ndarray, nxo: int, nyo: int, nzo: int, istro3: int, stro3: float):
        """Restart regridding of one field (interp.F90): old is C (nzo-1, nyo-1, nxo) as the
        continuation file holds it, i.e. Fortran (nxo, nyo-1, nzo-1)."""
        old = np.ascontiguousarray(old, dtype=np.float64)
        if old.shape != (nzo - 1, nyo - 1, nxo):
            raise ValueError(f"old field shape {old.shape} is not {(nzo - 1, nyo - 1, nxo)}")
        self._check(self.L.afid_interp_field(self.h, FIELD_IDS[name], _dp(old), nxo, nyo, nzo, istro3, stro3),
                    f"interp {name}")

    # ---- diagnostics -------------------------------------------------------------
    def CalcMaxCFL(self) -> float:
        v = ctypes.c_double()
        self._check(self.L.afid_calc_max_cfl(self.h, ctypes.byref(v)), "CalcMaxCFL")
        return v.value

    def CheckDivergence(self) -> float:
        v = ctypes.c_double()
        self._check(self.L.afid_check_divergence(self.h, ctypes.byref(v)), "CheckDivergence")
        return v.value

    def max_divergence(self) -> float:
        """CheckDivergence when every field is finite, else inf.  The maximum in
        CheckDivergence (CheckDivergence.F90:27-40) skips NaN as the
        reference's comparison does, so a cell that blew up would report a small
        value; GlobalQuantities' sums carry any NaN or inf through."""
        if not np.all(np.isfinite(self.GlobalQuantities())):
            return float("inf")
        return self.CheckDivergence()

    def CalcPlateNu(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        self._check(self.L.afid_calc_plate_nu(self.h, ctypes.byref(a), ctypes.byref(b)), "CalcPlateNu")
        return a.value, b.value

    def GlobalQuantities(self) -> np.ndarray:
        out = np.zeros(11)
        self._check(self.L.afid_global_quantities(self.h, _dp(out)), "GlobalQuantities")
        return out

    def CalcDissipationNu(self):
        mu, th = ctypes.c_double(), ctypes.c_double()
        de, dt_ = np.zeros(self.nxm), np.zeros(self.nxm)
        self._check(self.L.afid_calc_dissipation_nu(self.h, ctypes.byref(mu), ctypes.byref(th), _dp(de),
                                                    _dp(dt_)), "CalcDissipationNu")
        return mu.value, th.value, de, dt_

    # ---- profiling ----------------------------------------------------------------
    def profile(self, on: bool = True):
        self._check(self.L.afid_profile_enable(self.h, 1 if on else 0), "profile_enable")

    def profile_reset(self):
        self._check(self.L.afid_profile_reset(self.h), "profile_reset")

    def profile_read(self) -> dict:
        cap = 64
        buf = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * cap)()
        cnt = (ctypes.c_int64 * cap)()
        n = self.L.afid_profile_read(self.h, buf, 4096, ms, cnt, cap)
        if n < 0:
            self._check(n, "profile_read")
        names = buf.raw.split(b"\0")[:n]
        return {nm.decode(): (ms[q], cnt[q]) for q, nm in enumerate(names)}

"""ctypes wrapper of the CPU oracle (oracle/libdiablo_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker that the HIP path is compared with, and
the "port" CPU baseline in bench.py.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module.  The product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from diablo_amd.case import MAX_TH, ChannelConfig, grid_gy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libdiablo_oracle.so")

I32x = C.c_int32 * MAX_TH
F64x = C.c_double * MAX_TH


class _Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_th", C.c_int32),
        ("f_type", C.c_int32),
        ("u_bc_ymin", C.c_int32), ("v_bc_ymin", C.c_int32), ("w_bc_ymin", C.c_int32),
        ("u_bc_ymax", C.c_int32), ("v_bc_ymax", C.c_int32), ("w_bc_ymax", C.c_int32),
        ("th_bc_ymin", I32x), ("th_bc_ymax", I32x),
        ("nu", C.c_double), ("beta", C.c_double), ("lx", C.c_double), ("ly", C.c_double),
        ("lz", C.c_double), ("nu_v_scale", C.c_double), ("delta_t", C.c_double),
        ("i_ro", C.c_double), ("grav_x", C.c_double), ("grav_y", C.c_double), ("grav_z", C.c_double),
        ("px0", C.c_double), ("omega0", C.c_double), ("amp_omega0", C.c_double),
        ("u_bc_ymin_c1", C.c_double), ("v_bc_ymin_c1", C.c_double), ("w_bc_ymin_c1", C.c_double),
        ("u_bc_ymax_c1", C.c_double), ("v_bc_ymax_c1", C.c_double), ("w_bc_ymax_c1", C.c_double),
        ("th_bc_ymin_c1", F64x), ("th_bc_ymax_c1", F64x),
        ("ri", F64x), ("pr", F64x), ("drhodx", F64x), ("drhodz", F64x),
    ]


def build(quiet: bool = True, plant: int = 0, out: str | None = None) -> str:
    """Compiles oracle/libdiablo_oracle.so with the committed Makefile; with
    plant = n >= 1, a copy at `out` carrying planted mistake n (the tests of
    the pins use these copies; nothing else loads them)."""
    cmd = ["make", "-C", _HERE]
    if plant:
        if not out:
            raise ValueError("a planted build needs its own output path")
        cmd += [f"LIB={out}", f"PLANT={int(plant)}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    return out if plant else _LIB


_libs: dict = {}


def lib(path: str | None = None):
    """The oracle library (default: oracle/libdiablo_oracle.so, built if
    missing), loaded once per path with its signatures declared."""
    path = path or _LIB
    L = _libs.get(path)
    if L is None:
        if path == _LIB and not os.path.exists(_LIB):
            build()
        L = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        dp = C.POINTER(C.c_double)
        L.dor_create.restype = vp
        L.dor_create.argtypes = [C.POINTER(_Params), dp]
        L.dor_destroy.argtypes = [vp]
        L.dor_field.restype = dp
        L.dor_field.argtypes = [vp, i]
        L.dor_vec.restype = dp
        L.dor_vec.argtypes = [vp, i]
        L.dor_info.restype = i
        L.dor_info.argtypes = [vp, i]
        L.dor_set_time.argtypes = [vp, d]
        L.dor_set_dt.argtypes = [vp, d]
        L.dor_get_dt.restype = d
        L.dor_get_dt.argtypes = [vp]
        L.dor_courant.restype = d
        L.dor_courant.argtypes = [vp, d]
        L.dor_courant_request.argtypes = [vp, d]
        L.dor_flow_to_file.argtypes = [vp, i, dp]
        L.dor_flow_from_file.argtypes = [vp, i, dp]
        L.dor_set_threads.argtypes = [i]
        for name in ("dor_fft_xz_to_fourier", "dor_fft_xz_to_physical"):
            getattr(L, name).argtypes = [vp, dp, i, i]
        L.dor_rk_chan_1.argtypes = [vp, i]
        for name in ("dor_rem_div_chan", "dor_poisson_p_chan", "dor_apply_bc_vel", "dor_apply_bc_scalar"):
            getattr(L, name).argtypes = [vp]
        L.dor_thomas_real.argtypes = [dp, dp, dp, dp, i, i]
        L.dor_thomas_complex.argtypes = [dp, dp, dp, dp, i, i]
        L.dor_fft_c2c.argtypes = [dp, i, i]
        L.dor_r2c.argtypes = [dp, dp, i]
        L.dor_c2r.argtypes = [dp, dp, i]
        L.dor_stats_chan.argtypes = [vp, dp, dp]
        L.dor_tke_chan.argtypes = [vp, dp]
        L.dor_tke_les.argtypes = [vp, dp]
        L.dor_les_chan.argtypes = [vp, dp]
        L.dor_set_les.argtypes = [vp, i, i]
        _libs[path] = L
    return L


def _dp(a: np.ndarray):
    assert a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def params_of(cfg: ChannelConfig) -> _Params:
    p = _Params()
    for name, _ in _Params._fields_:
        v = getattr(cfg, name)
        if isinstance(v, list):
            arr = getattr(p, name)
            for n, x in enumerate(v[:MAX_TH]):
                arr[n] = x
        else:
            setattr(p, name, v)
    return p


# SAVE_STATS_CHAN statistics, in the order of diablo_oracle.h's DOR_ST_*
STATS = ("ume", "vme", "wme", "urms", "vrms", "wrms", "uv", "uw", "wv", "dudy", "dwdy", "cp", "shear",
         "omega_x", "omega_y", "omega_z", "thme", "thrms", "thv", "dthdy", "pe_diss")

# field ids (diablo_oracle.h)
U1, U2, U3, P, R1, R2, R3, F1, F2, F3, S1 = range(11)
TH0, FTH0, RTH0, CTH0, CFTH0, CRTH0 = 16, 24, 32, 40, 48, 56
NUT, KAPPA0 = 11, 20
GY, GYF, DY, DYF, KX, KZ, KX2, KZ2, GX, GZ = range(10)


class Oracle:
    """The reference's channel state and call surface on the CPU."""

    def __init__(self, cfg: ChannelConfig, gy_1: np.ndarray | None = None, threads: int = 0,
                 library: str | None = None):
        cfg.validate()
        self.cfg = cfg
        self.L = lib(library)
        if threads:
            self.L.dor_set_threads(threads)
        gy = np.ascontiguousarray(grid_gy(cfg) if gy_1 is None else gy_1, dtype=np.float64)
        self._p = params_of(cfg)
        self.h = self.L.dor_create(C.byref(self._p), _dp(gy))
        if not self.h:
            raise ValueError("oracle rejected the configuration")
        self.time = 0.0

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.L.dor_destroy(h)
            self.h = None

    def field(self, fid: int) -> np.ndarray:
        ptr = self.L.dor_field(self.h, fid)
        shp = self.cfg.field_shape()
        return np.ctypeslib.as_array(ptr, shape=shp)

    def vec(self, vid: int) -> np.ndarray:
        n = {KX: self.cfg.nxp + 1, KX2: self.cfg.nxp + 1, KZ: self.cfg.tnkz + 1, KZ2: self.cfg.tnkz + 1,
             GX: self.cfg.nx + 2, GZ: self.cfg.nz + 2}.get(vid, self.cfg.ny + 2)
        return np.ctypeslib.as_array(self.L.dor_vec(self.h, vid), shape=(n,)).copy()

    def info(self, what: int) -> int:
        return self.L.dor_info(self.h, what)

    # -- reference call surface --
    def fft_xz_to_fourier(self, fid: int, jmin: int, jmax: int):
        self.L.dor_fft_xz_to_fourier(self.h, self.L.dor_field(self.h, fid), jmin, jmax)

    def fft_xz_to_physical(self, fid: int, jmin: int, jmax: int):
        self.L.dor_fft_xz_to_physical(self.h, self.L.dor_field(self.h, fid), jmin, jmax)

    def fft_array_to_fourier(self, a: np.ndarray, jmin: int, jmax: int):
        self.L.dor_fft_xz_to_fourier(self.h, _dp(a), jmin, jmax)

    def fft_array_to_physical(self, a: np.ndarray, jmin: int, jmax: int):
        self.L.dor_fft_xz_to_physical(self.h, _dp(a), jmin, jmax)

    def rk_chan_1(self, rk_step: int, time: float | None = None):
        """RK_CHAN_1 with TIME = time (default: the current step's start time)."""
        self.L.dor_set_time(self.h, self.time if time is None else time)
        self.L.dor_rk_chan_1(self.h, rk_step)

    def step(self, courant_cfl: float | None = None):
        """One time step: RK_STEP = 1..3 then TIME += DELTA_T (for/diablo.f:70-109).
        courant_cfl: run COURANT inside RK_STEP 3 with that CFL (VARIABLE_DT on a
        step with MOD(TIME_STEP, UPDATE_DT) = 0); TIME then advances by the new
        DELTA_T, as the reference's does."""
        for k in (1, 2, 3):
            if k == 3 and courant_cfl is not None:
                self.L.dor_courant_request(self.h, courant_cfl)
            self.rk_chan_1(k)
        self.time += self.delta_t

    @property
    def delta_t(self) -> float:
        return float(self.L.dor_get_dt(self.h))

    def set_delta_t(self, dt: float):
        self.L.dor_set_dt(self.h, dt)

    def set_les(self, les: bool, model: bool = True):
        """RK_CHAN_1's LES switches: les = the eddy viscosity in the
        Crank-Nicolson terms; model = les_chan called in each substep."""
        self.L.dor_set_les(self.h, int(bool(les)), int(bool(model)))

    def courant(self, cfl: float) -> float:
        """COURANT on the physical U1..U3 arrays as they stand; returns DELTA_T."""
        return float(self.L.dor_courant(self.h, cfl))

    def rem_div_chan(self):
        self.L.dor_rem_div_chan(self.h)

    def poisson_p_chan(self):
        self.L.dor_poisson_p_chan(self.h)

    def apply_bc_vel(self):
        self.L.dor_apply_bc_vel(self.h)

    def apply_bc_scalar(self):
        self.L.dor_apply_bc_scalar(self.h)

    def stats_chan(self) -> tuple[dict, float]:
        """SAVE_STATS_CHAN's statistics: {name: profile at J = 1..NY}, UBULK."""
        ny = self.cfg.ny
        out = np.zeros(len(STATS) * ny)
        ub = np.zeros(1)
        self.L.dor_stats_chan(self.h, _dp(out), _dp(ub))
        return {k: out[q * ny:(q + 1) * ny].copy() for q, k in enumerate(STATS)}, float(ub[0])

    def tke_chan(self) -> np.ndarray:
        """tkebudget_chan's epsilon at J = 1..NY (entry 0 depends on the S1
        scratch's previous contents, as in the reference)."""
        out = np.zeros(self.cfg.ny)
        self.L.dor_tke_chan(self.h, _dp(out))
        return out

    def tke_les(self) -> np.ndarray:
        """tkebudget_chan_les's epsilon_sgs at J = 1..NY (LES on)."""
        out = np.zeros(self.cfg.ny)
        self.L.dor_tke_les(self.h, _dp(out))
        return out

    def les_chan(self) -> np.ndarray:
        """les_chan (constant Smagorinsky): updates CF1-3 (and the velocity
        BC rows); returns NU_T in field layout."""
        out = np.zeros(self.cfg.field_shape())
        self.L.dor_les_chan(self.h, _dp(out))
        return out

    # -- restart-file layout (WriteHDF5 / ReadHDF5) --
    def flow_to_file(self, fid: int) -> np.ndarray:
        """The array WriteHDF5 writes for field fid, as (NZ, NY, NX) C order
        (= tmp(1:NX, 1:NY, 1:NZ) in Fortran order)."""
        out = np.zeros((self.cfg.nz, self.cfg.ny, self.cfg.nx))
        self.L.dor_flow_to_file(self.h, fid, _dp(out))
        return out

    def flow_from_file(self, fid: int, a: np.ndarray):
        """ReadHDF5's conversion of a file array into field fid."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.cfg.nz, self.cfg.ny, self.cfg.nx)
        self.L.dor_flow_from_file(self.h, fid, _dp(a))

    # -- initial state: the CREATE_FLOW_CHAN / CREATE_TH_CHAN sequence --
    def initialize(self, u1, u2, u3, ths):
        """Physical start fields -> Fourier state (for/set_ics.f:110-168, 251-253)."""
        cfg = self.cfg
        self.field(U1)[...] = u1
        self.field(U2)[...] = u2
        self.field(U3)[...] = u3
        self.field(P)[...] = 0.0
        for n in range(cfg.n_th):
            s1 = self.field(S1)
            s1[...] = ths[n]
            self.fft_xz_to_fourier(S1, 0, cfg.ny + 1)
            self.field(CTH0 + n)[...] = s1
        for f in (U1, U2, U3, P):
            self.fft_xz_to_fourier(f, 0, cfg.ny + 1)
        self.apply_bc_vel()
        self.rem_div_chan()
        self.time = 0.0

    def state(self) -> dict:
        """Copies of the Fourier state arrays (Fortran layout, float64)."""
        d = {"CU1": self.field(U1).copy(), "CU2": self.field(U2).copy(), "CU3": self.field(U3).copy(),
             "CP": self.field(P).copy(), "CF1": self.field(F1).copy(), "CF2": self.field(F2).copy(),
             "CF3": self.field(F3).copy()}
        for n in range(self.cfg.n_th):
            d[f"CTH{n}"] = self.field(CTH0 + n).copy()
            d[f"CFTH{n}"] = self.field(CFTH0 + n).copy()
        return d

    def load_state(self, d: dict):
        self.field(U1)[...] = d["CU1"]
        self.field(U2)[...] = d["CU2"]
        self.field(U3)[...] = d["CU3"]
        self.field(P)[...] = d["CP"]
        for k, f in (("CF1", F1), ("CF2", F2), ("CF3", F3)):
            if k in d:
                self.field(f)[...] = d[k]
        for n in range(self.cfg.n_th):
            self.field(CTH0 + n)[...] = d[f"CTH{n}"]
            if f"CFTH{n}" in d:
                self.field(CFTH0 + n)[...] = d[f"CFTH{n}"]


def thomas_real(a, b, c, g):
    """In-place THOMAS_*_REAL_MPI on (NY+2, NX) arrays (row j, column i)."""
    ny = a.shape[0] - 2
    nx = a.shape[1]
    lib().dor_thomas_real(_dp(a), _dp(b), _dp(c), _dp(g), ny, nx)


def thomas_complex(a, b, c, g):
    """In-place THOMAS_*_COMPLEX_MPI on (NY+2, NXP+1) arrays; the last column is
    the padding column and is not solved."""
    ny = a.shape[0] - 2
    nxp = a.shape[1] - 1
    lib().dor_thomas_complex(_dp(a), _dp(b), _dp(c), _dp(g), ny, nxp)


def fft_c2c(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().dor_fft_c2c(_dp(y), y.size, sign)
    return y


def r2c(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.size // 2 + 1, dtype=np.complex128)
    lib().dor_r2c(_dp(x), _dp(out), x.size)
    return out


def c2r(X: np.ndarray, n: int) -> np.ndarray:
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.zeros(n)
    lib().dor_c2r(_dp(X), _dp(out), n)
    return out

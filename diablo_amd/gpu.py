"""Python host mirror of the reference call surface over libdiablo_gpu.so.

``GpuChannel`` exposes the reference's subroutine names (RK_CHAN_1,
REM_DIV_CHAN, FFT_XZ_TO_FOURIER, FFT_XZ_TO_PHYSICAL, APPLY_BC_VEL) with the
same argument meaning, and goes through the C-ABI declared in
include/diablo_gpu.h.  There is no CPU fallback: if the HIP library is
missing or cannot create a context, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .case import ChannelConfig, grid_gy, rkw3_table

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdiablo_gpu.so")

# include/diablo_gpu.h: DGPU_ABI_VERSION, DGPU_MAX_TH (active scalars the device path carries)
ABI_VERSION = 2
DGPU_MAX_TH = 1
I32x = C.c_int32 * DGPU_MAX_TH
F64x = C.c_double * DGPU_MAX_TH
DP = C.POINTER(C.c_double)

# field ids (include/diablo_gpu.h)
CU1, CU2, CU3, CP, CF1, CF2, CF3 = 0, 1, 2, 3, 4, 5, 6
CTH0, CFTH0 = 16, 24
KERNEL_KINDS = ("zinv", "xfused", "zfwd_nl", "cn", "remdiv")


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("n_th", C.c_int32),
        ("precision", C.c_int32), ("f_type", C.c_int32),
        ("u_bc_ymin", C.c_int32), ("v_bc_ymin", C.c_int32), ("w_bc_ymin", C.c_int32),
        ("u_bc_ymax", C.c_int32), ("v_bc_ymax", C.c_int32), ("w_bc_ymax", C.c_int32),
        ("th_bc_ymin", I32x), ("th_bc_ymax", I32x),
        ("nu", C.c_double), ("beta", C.c_double), ("lx", C.c_double), ("ly", C.c_double),
        ("lz", C.c_double), ("nu_v_scale", C.c_double),
        ("i_ro", C.c_double), ("grav_x", C.c_double), ("grav_y", C.c_double), ("grav_z", C.c_double),
        ("px0", C.c_double), ("omega0", C.c_double), ("amp_omega0", C.c_double),
        ("u_bc_ymin_c1", C.c_double), ("v_bc_ymin_c1", C.c_double), ("w_bc_ymin_c1", C.c_double),
        ("u_bc_ymax_c1", C.c_double), ("v_bc_ymax_c1", C.c_double), ("w_bc_ymax_c1", C.c_double),
        ("th_bc_ymin_c1", F64x), ("th_bc_ymax_c1", F64x),
        ("ri", F64x), ("pr", F64x), ("drhodx", F64x), ("drhodz", F64x),
        ("gy", DP), ("gyf", DP), ("dy", DP), ("dyf", DP),
        ("kx", DP), ("kz", DP), ("kx2", DP), ("kz2", DP),
    ]


EXPORTS = (
    "dgpu_abi_version", "dgpu_build_info", "dgpu_create", "dgpu_destroy", "dgpu_last_error",
    "dgpu_set_field", "dgpu_get_field", "dgpu_rk_chan_1", "dgpu_rem_div_chan", "dgpu_poisson_p_chan",
    "dgpu_apply_bc_vel",
    "dgpu_fft_xz_to_fourier", "dgpu_fft_xz_to_physical", "dgpu_synchronize",
    "dgpu_set_kernel_timing", "dgpu_kernel_times", "dgpu_set_graphs",
    "dgpu_ensemble_create", "dgpu_ensemble_destroy", "dgpu_ensemble_size", "dgpu_ensemble_member",
    "dgpu_ensemble_last_error", "dgpu_ensemble_set_graphs", "dgpu_ensemble_rk_chan_1",
    "dgpu_ensemble_synchronize", "dgpu_stats_chan", "dgpu_courant_request", "dgpu_courant",
    "dgpu_get_flow_physical", "dgpu_set_flow_physical", "dgpu_tke_chan", "dgpu_les_chan", "dgpu_set_les", "dgpu_set_zdft", "dgpu_autotune", "dgpu_tke_les",
    "dgpu_stats_chan_noalloc", "dgpu_set_column_shape", "dgpu_get_column_shape", "dgpu_ensemble_autotune",
)

# dgpu_stats_chan profiles, in the order of the header's DGPU_ST_*
STATS = ("ume", "vme", "wme", "urms", "vrms", "wrms", "uv", "uw", "wv", "dudy", "dwdy", "cp", "shear",
         "omega_x", "omega_y", "omega_z", "thme", "thrms", "thv", "dthdy", "pe_diss")

_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libdiablo_gpu.so and declares its signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m diablo_amd.build`")
    L = C.CDLL(path)
    vp, i32 = C.c_void_p, C.c_int32
    L.dgpu_abi_version.restype = i32
    if L.dgpu_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path}: ABI version {L.dgpu_abi_version()}, expected {ABI_VERSION} (rebuild it)")
    L.dgpu_build_info.restype = C.c_char_p
    L.dgpu_create.restype = i32
    L.dgpu_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.dgpu_destroy.restype = i32
    L.dgpu_destroy.argtypes = [vp]
    L.dgpu_last_error.restype = C.c_char_p
    L.dgpu_last_error.argtypes = [vp]
    for name in ("dgpu_set_field", "dgpu_get_field"):
        getattr(L, name).restype = i32
        getattr(L, name).argtypes = [vp, i32, DP]
    L.dgpu_rk_chan_1.restype = i32
    L.dgpu_rk_chan_1.argtypes = [vp, i32, C.c_double, DP, DP, DP]
    for name in ("dgpu_rem_div_chan", "dgpu_poisson_p_chan", "dgpu_apply_bc_vel", "dgpu_synchronize"):
        getattr(L, name).restype = i32
        getattr(L, name).argtypes = [vp]
    for name in ("dgpu_fft_xz_to_fourier", "dgpu_fft_xz_to_physical"):
        getattr(L, name).restype = i32
        getattr(L, name).argtypes = [vp, DP, i32, i32]
    L.dgpu_ensemble_create.restype = i32
    L.dgpu_ensemble_create.argtypes = [C.POINTER(Params), i32, C.POINTER(vp)]
    L.dgpu_ensemble_destroy.restype = i32
    L.dgpu_ensemble_destroy.argtypes = [vp]
    L.dgpu_ensemble_size.restype = i32
    L.dgpu_ensemble_size.argtypes = [vp]
    L.dgpu_ensemble_member.restype = vp
    L.dgpu_ensemble_member.argtypes = [vp, i32]
    L.dgpu_ensemble_last_error.restype = C.c_char_p
    L.dgpu_ensemble_last_error.argtypes = [vp]
    L.dgpu_ensemble_set_graphs.restype = i32
    L.dgpu_ensemble_set_graphs.argtypes = [vp, i32]
    L.dgpu_ensemble_rk_chan_1.restype = i32
    L.dgpu_ensemble_rk_chan_1.argtypes = [vp, i32, DP, DP, DP]
    L.dgpu_ensemble_synchronize.restype = i32
    L.dgpu_ensemble_synchronize.argtypes = [vp]
    L.dgpu_set_graphs.restype = i32
    L.dgpu_set_graphs.argtypes = [vp, i32]
    L.dgpu_set_kernel_timing.restype = i32
    L.dgpu_set_kernel_timing.argtypes = [vp, i32]
    L.dgpu_kernel_times.restype = i32
    L.dgpu_kernel_times.argtypes = [vp, DP, C.POINTER(C.c_int64), i32]
    for name in ("dgpu_stats_chan", "dgpu_stats_chan_noalloc"):
        getattr(L, name).restype = i32
        getattr(L, name).argtypes = [vp, DP, DP]
    L.dgpu_tke_chan.restype = i32
    L.dgpu_tke_chan.argtypes = [vp, DP]
    L.dgpu_les_chan.restype = i32
    L.dgpu_les_chan.argtypes = [vp, DP]
    L.dgpu_set_les.restype = i32
    L.dgpu_set_les.argtypes = [vp, i32, i32]
    L.dgpu_tke_les.restype = i32
    L.dgpu_tke_les.argtypes = [vp, DP]
    L.dgpu_set_zdft.restype = i32
    L.dgpu_set_zdft.argtypes = [vp, i32]
    L.dgpu_set_column_shape.restype = i32
    L.dgpu_set_column_shape.argtypes = [vp, i32, i32]
    L.dgpu_get_column_shape.restype = i32
    L.dgpu_get_column_shape.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    L.dgpu_ensemble_autotune.restype = i32
    L.dgpu_ensemble_autotune.argtypes = [vp, i32, DP, DP, DP, C.POINTER(i32), C.POINTER(i32), DP]
    L.dgpu_autotune.restype = i32
    L.dgpu_autotune.argtypes = [vp, i32, C.c_double, C.POINTER(i32), C.POINTER(i32), DP]
    L.dgpu_get_flow_physical.restype = i32
    L.dgpu_get_flow_physical.argtypes = [vp, i32, DP]
    L.dgpu_set_flow_physical.restype = i32
    L.dgpu_set_flow_physical.argtypes = [vp, i32, DP]
    L.dgpu_courant_request.restype = i32
    L.dgpu_courant_request.argtypes = [vp]
    L.dgpu_courant.restype = i32
    L.dgpu_courant.argtypes = [vp, C.c_double, C.POINTER(C.c_double), DP]
    _lib = L
    return L


def _dp(a: np.ndarray):
    if not (a.dtype in (np.float64, np.complex128) and a.flags.c_contiguous):
        raise ValueError("expected a C-contiguous float64 array in Fortran field layout")
    return a.ctypes.data_as(DP)


PRECISIONS = {"fp64": 0, "mixed": 1, "mixed_f64state": 2}


def params_of(cfg: ChannelConfig, gy_1: np.ndarray, precision: str = "fp64") -> Params:
    if cfg.n_th > DGPU_MAX_TH:
        raise DgpuError(f"N_TH = {cfg.n_th}: the device path carries at most {DGPU_MAX_TH} scalar")
    p = Params()
    for name, _ in Params._fields_:
        if name in ("gy", "gyf", "dy", "dyf", "kx", "kz", "kx2", "kz2", "precision"):
            continue
        v = getattr(cfg, name)
        if isinstance(v, list):
            arr = getattr(p, name)
            for n, x in enumerate(v[:DGPU_MAX_TH]):
                arr[n] = x
        else:
            setattr(p, name, v)
    p.precision = PRECISIONS[precision]
    p.gy = _dp(gy_1)
    return p


class DgpuError(RuntimeError):
    pass


class GpuChannel:
    """The channel state on one MI355X, driven through the C-ABI."""

    def __init__(self, cfg: ChannelConfig, gy_1: np.ndarray | None = None, precision: str = "fp64"):
        """precision: "fp64" (every kernel FP64), "mixed" (FP32 transforms,
        products, nonlinear terms and column solves on an FP32 working copy of
        the state; the low modes I <= 2, |kz index| <= 2, among them the (0,0)
        mean mode and the pressure gauge, solved in FP64) or "mixed_f64state"
        (FP32 transforms, products and nonlinear terms; FP64 state and column
        solves)."""
        cfg.validate()
        self.cfg = cfg
        self.precision = precision
        self.L = load_library()
        self._gy = np.ascontiguousarray(grid_gy(cfg) if gy_1 is None else gy_1, dtype=np.float64)
        self._p = params_of(cfg, self._gy, precision)
        h = C.c_void_p()
        rc = self.L.dgpu_create(C.byref(self._p), C.byref(h))
        if rc != 0 or not h.value:
            raise DgpuError(f"dgpu_create failed ({rc}) for {cfg.nx}x{cfg.ny}x{cfg.nz}")
        self.h = h
        self.time = 0.0
        self.h_bar, self.beta_bar, self.zeta_bar = (np.ascontiguousarray(a) for a in rkw3_table(cfg.delta_t))

    @classmethod
    def attach(cls, cfg: ChannelConfig, handle: int, precision: str, gy_1: np.ndarray) -> "GpuChannel":
        """Wraps a context owned by someone else (an ensemble member)."""
        self = cls.__new__(cls)
        self.cfg = cfg
        self.precision = precision
        self.L = load_library()
        self._gy = gy_1
        self.h = C.c_void_p(handle)
        self._owned = False
        self.time = 0.0
        self.h_bar, self.beta_bar, self.zeta_bar = (np.ascontiguousarray(a) for a in rkw3_table(cfg.delta_t))
        return self

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value and getattr(self, "_owned", True):
            self.L.dgpu_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int, what: str):
        if rc != 0:
            msg = self.L.dgpu_last_error(self.h)
            raise DgpuError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    # -- state exchange (Fortran-layout host arrays) --
    def set_field(self, fid: int, a: np.ndarray):
        self._chk(self.L.dgpu_set_field(self.h, fid, _dp(a)), "dgpu_set_field")

    def get_field(self, fid: int) -> np.ndarray:
        a = np.zeros(self.cfg.field_shape())
        self._chk(self.L.dgpu_get_field(self.h, fid, _dp(a)), "dgpu_get_field")
        return a

    def state(self) -> dict:
        d = {"CU1": self.get_field(CU1), "CU2": self.get_field(CU2), "CU3": self.get_field(CU3),
             "CP": self.get_field(CP), "CF1": self.get_field(CF1), "CF2": self.get_field(CF2),
             "CF3": self.get_field(CF3)}
        for n in range(self.cfg.n_th):
            d[f"CTH{n}"] = self.get_field(CTH0 + n)
            d[f"CFTH{n}"] = self.get_field(CFTH0 + n)
        return d

    def load_state(self, d: dict):
        for k, f in (("CU1", CU1), ("CU2", CU2), ("CU3", CU3), ("CP", CP), ("CF1", CF1), ("CF2", CF2),
                     ("CF3", CF3)):
            if k in d:
                self.set_field(f, np.ascontiguousarray(d[k]))
        for n in range(self.cfg.n_th):
            self.set_field(CTH0 + n, np.ascontiguousarray(d[f"CTH{n}"]))
            if f"CFTH{n}" in d:
                self.set_field(CFTH0 + n, np.ascontiguousarray(d[f"CFTH{n}"]))

    # -- the reference call surface --
    def rk_chan_1(self, rk_step: int, time: float | None = None):
        t = self.time if time is None else time
        self._chk(self.L.dgpu_rk_chan_1(self.h, rk_step, t, _dp(self.h_bar), _dp(self.beta_bar),
                                        _dp(self.zeta_bar)), "dgpu_rk_chan_1")

    def step(self, n: int = 1, courant_cfl: float | None = None):
        """n time steps: RK_STEP = 1..3 each, then TIME += DELTA_T (for/diablo.f:70-109).
        courant_cfl: VARIABLE_DT with UPDATE_DT = 1 -- COURANT inside every RK_STEP 3
        with that CFL; the new DELTA_T applies from the next step, and TIME advances
        by it, as the reference's does."""
        for _ in range(n):
            for k in (1, 2, 3):
                if k == 3 and courant_cfl is not None:
                    self._chk(self.L.dgpu_courant_request(self.h), "dgpu_courant_request")
                self.rk_chan_1(k)
            if courant_cfl is not None:
                self.courant(courant_cfl)
            self.time += self.delta_t

    @property
    def delta_t(self) -> float:
        return getattr(self, "_dt", self.cfg.delta_t)

    def set_delta_t(self, dt: float):
        self._dt = float(dt)
        self.h_bar = np.ascontiguousarray(rkw3_table(dt)[0])

    def courant(self, cfl: float) -> float:
        """DELTA_T from the maxima the last RK_STEP 3 recorded; updates H_BAR."""
        dt = C.c_double()
        hb = np.zeros(3)
        self._chk(self.L.dgpu_courant(self.h, cfl, C.byref(dt), _dp(hb)), "dgpu_courant")
        self._dt = dt.value
        self.h_bar = hb
        return dt.value

    def rem_div_chan(self):
        self._chk(self.L.dgpu_rem_div_chan(self.h), "dgpu_rem_div_chan")

    def poisson_p_chan(self):
        self._chk(self.L.dgpu_poisson_p_chan(self.h), "dgpu_poisson_p_chan")

    def apply_bc_vel(self):
        self._chk(self.L.dgpu_apply_bc_vel(self.h), "dgpu_apply_bc_vel")

    def tke_chan(self) -> np.ndarray:
        """tkebudget_chan's epsilon at J = 1..NY from the device state."""
        out = np.zeros(self.cfg.ny)
        self._chk(self.L.dgpu_tke_chan(self.h, _dp(out)), "dgpu_tke_chan")
        return out

    def les_chan(self) -> np.ndarray:
        """les_chan's constant Smagorinsky stress divergence added to CF1-3
        (and the velocity BC rows set); returns NU_T in field layout."""
        out = np.zeros(self.cfg.field_shape())
        self._chk(self.L.dgpu_les_chan(self.h, _dp(out)), "dgpu_les_chan")
        return out

    def set_zdft(self, on: bool):
        """Mixed precision: z transforms as split-BF16 DFT-as-GEMM on the
        matrix cores (NZ <= 256)."""
        self._chk(self.L.dgpu_set_zdft(self.h, int(bool(on))), "dgpu_set_zdft")

    def autotune(self, steps: int = 3) -> dict:
        """Times the launch-shape candidates on this context's state, keeps
        the fastest; the state is left as it was."""
        xc, zd, ms = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        self._chk(self.L.dgpu_autotune(self.h, steps, self.delta_t, C.byref(xc), C.byref(zd), C.byref(ms)),
                  "dgpu_autotune")
        return {"xchunk": xc.value, "zdft": bool(zd.value), "column_waves": self.column_shape(),
                "column_fused": self.column_fused(), "ms_per_step": ms.value}

    def set_column_shape(self, waves: int, fused: int = -1):
        """Waves per block of the segmented column kernels (4, 8, 16), or 0
        for the whole-column sweeps; fused (FP32 state or FP64): 1 the
        one-kernel column stage, 0 two kernels, -1 the library's choice by
        NY (and, in FP64, the waves)."""
        self._chk(self.L.dgpu_set_column_shape(self.h, int(waves), int(fused)), "dgpu_set_column_shape")

    def _column(self):
        w, f = C.c_int32(0), C.c_int32(0)
        self._chk(self.L.dgpu_get_column_shape(self.h, C.byref(w), C.byref(f)), "dgpu_get_column_shape")
        return w.value, bool(f.value)

    def column_shape(self) -> int:
        return self._column()[0]

    def column_fused(self) -> bool:
        return self._column()[1]

    def tke_les(self) -> np.ndarray:
        """tkebudget_chan_les's epsilon_sgs at J = 1..NY (LES on; J = 1 is NaN
        as in the reference)."""
        out = np.zeros(self.cfg.ny)
        self._chk(self.L.dgpu_tke_les(self.h, _dp(out)), "dgpu_tke_les")
        return out

    def set_les(self, les: bool, model: bool = True):
        """RK_CHAN_1's LES switches: les = the eddy viscosity in the
        Crank-Nicolson terms (physical-space solves); model = les_chan called
        in each substep."""
        self._chk(self.L.dgpu_set_les(self.h, int(bool(les)), int(bool(model))), "dgpu_set_les")

    def get_flow_physical(self, fid: int) -> np.ndarray:
        """WriteHDF5's array for field fid (CU1..CU3, CP, CTH0+n): (NZ, NY, NX)
        C order = tmp(1:NX, 1:NY, 1:NZ) in Fortran order."""
        out = np.zeros((self.cfg.nz, self.cfg.ny, self.cfg.nx))
        self._chk(self.L.dgpu_get_flow_physical(self.h, fid, _dp(out)), "dgpu_get_flow_physical")
        return out

    def set_flow_physical(self, fid: int, a: np.ndarray):
        """ReadHDF5's conversion of a file array into field fid."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.cfg.nz, self.cfg.ny, self.cfg.nx):
            raise ValueError(f"file array must be (NZ, NY, NX), got {a.shape}")
        self._chk(self.L.dgpu_set_flow_physical(self.h, fid, _dp(a)), "dgpu_set_flow_physical")

    def stats_chan(self) -> tuple[dict, float]:
        """SAVE_STATS_CHAN's statistics from the device state:
        {name: profile at J = 1..NY}, UBULK."""
        ny = self.cfg.ny
        out = np.zeros(len(STATS) * ny)
        ub = np.zeros(1)
        self._chk(self.L.dgpu_stats_chan(self.h, _dp(out), _dp(ub)), "dgpu_stats_chan")
        return {k: out[q * ny:(q + 1) * ny].copy() for q, k in enumerate(STATS)}, float(ub[0])

    def stats_chan_noalloc(self) -> tuple[dict, float]:
        """stats_chan through dgpu_stats_chan_noalloc: the same statistics with
        the library's fixed device array instead of an allocated buffer."""
        ny = self.cfg.ny
        out = np.zeros(len(STATS) * ny)
        ub = np.zeros(1)
        self._chk(self.L.dgpu_stats_chan_noalloc(self.h, _dp(out), _dp(ub)), "dgpu_stats_chan_noalloc")
        return {k: out[q * ny:(q + 1) * ny].copy() for q, k in enumerate(STATS)}, float(ub[0])

    def fft_xz_to_fourier(self, a: np.ndarray, jmin: int, jmax: int):
        self._chk(self.L.dgpu_fft_xz_to_fourier(self.h, _dp(a), jmin, jmax), "dgpu_fft_xz_to_fourier")

    def fft_xz_to_physical(self, a: np.ndarray, jmin: int, jmax: int):
        self._chk(self.L.dgpu_fft_xz_to_physical(self.h, _dp(a), jmin, jmax), "dgpu_fft_xz_to_physical")

    def synchronize(self):
        self._chk(self.L.dgpu_synchronize(self.h), "dgpu_synchronize")

    def set_graphs(self, on: bool):
        self._chk(self.L.dgpu_set_graphs(self.h, 1 if on else 0), "dgpu_set_graphs")

    def set_kernel_timing(self, on: bool):
        self._chk(self.L.dgpu_set_kernel_timing(self.h, 1 if on else 0), "dgpu_set_kernel_timing")

    def kernel_times(self) -> dict:
        ms = (C.c_double * 5)()
        cnt = (C.c_int64 * 5)()
        self._chk(self.L.dgpu_kernel_times(self.h, ms, cnt, 5), "dgpu_kernel_times")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(KERNEL_KINDS)}

    # -- CREATE_FLOW_CHAN / CREATE_TH_CHAN on the device --
    def initialize(self, u1, u2, u3, ths):
        """Physical start fields -> device Fourier state, then APPLY_BC_VEL and
        REM_DIV_CHAN (for/set_ics.f:110-168, 251-253)."""
        cfg = self.cfg
        for fid, a in ((CU1, u1), (CU2, u2), (CU3, u3)):
            w = np.ascontiguousarray(a, dtype=np.float64).copy()
            self.fft_xz_to_fourier(w, 0, cfg.ny + 1)
            self.set_field(fid, w)
        self.set_field(CP, np.zeros(cfg.field_shape()))
        for n in range(cfg.n_th):
            w = np.ascontiguousarray(ths[n], dtype=np.float64).copy()
            self.fft_xz_to_fourier(w, 0, cfg.ny + 1)
            self.set_field(CTH0 + n, w)
        self.apply_bc_vel()
        self.rem_div_chan()
        self.synchronize()
        self.time = 0.0


class GpuEnsemble:
    """Independent channel runs (their own sizes, parameters and state)
    advanced together through dgpu_ensemble_rk_chan_1."""

    def __init__(self, cfgs, precision: str = "mixed"):
        self.cfgs = list(cfgs)
        self.L = load_library()
        n = len(self.cfgs)
        self._gys = [np.ascontiguousarray(grid_gy(c), dtype=np.float64) for c in self.cfgs]
        arr = (Params * n)()
        for i, c in enumerate(self.cfgs):
            c.validate()
            arr[i] = params_of(c, self._gys[i], precision)
        self._params = arr
        h = C.c_void_p()
        rc = self.L.dgpu_ensemble_create(arr, n, C.byref(h))
        if rc != 0 or not h.value:
            raise DgpuError(f"dgpu_ensemble_create failed ({rc}) for {n} members")
        self.h = h
        self.members = [GpuChannel.attach(c, self.L.dgpu_ensemble_member(h, i), precision, self._gys[i])
                        for i, c in enumerate(self.cfgs)]
        tabs = [rkw3_table(c.delta_t) for c in self.cfgs]
        self.h_bar = np.ascontiguousarray(np.concatenate([t[0] for t in tabs]))
        self.beta_bar = np.ascontiguousarray(np.concatenate([t[1] for t in tabs]))
        self.zeta_bar = np.ascontiguousarray(np.concatenate([t[2] for t in tabs]))

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.dgpu_ensemble_last_error(self.h)
            raise DgpuError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def set_graphs(self, on: bool):
        self._chk(self.L.dgpu_ensemble_set_graphs(self.h, 1 if on else 0), "dgpu_ensemble_set_graphs")

    def rk_chan_1(self, rk_step: int):
        self._chk(self.L.dgpu_ensemble_rk_chan_1(self.h, rk_step, _dp(self.h_bar), _dp(self.beta_bar),
                                                 _dp(self.zeta_bar)), "dgpu_ensemble_rk_chan_1")

    def step(self, n: int = 1):
        for _ in range(n):
            for k in (1, 2, 3):
                self.rk_chan_1(k)
            for m in self.members:
                m.time += m.cfg.delta_t

    def autotune(self, steps: int = 2) -> dict:
        """Times the batched launch shapes on the members' states (every
        member's x-kernel planes per block and column waves), keeps the
        fastest; the states are left as they were."""
        xc, w, ms = C.c_int32(0), C.c_int32(0), C.c_double(0.0)
        self._chk(self.L.dgpu_ensemble_autotune(self.h, steps, _dp(self.h_bar), _dp(self.beta_bar),
                                                _dp(self.zeta_bar), C.byref(xc), C.byref(w), C.byref(ms)),
                  "dgpu_ensemble_autotune")
        return {"xchunk": xc.value, "column_waves": w.value, "ms_per_step": ms.value}

    def synchronize(self):
        self._chk(self.L.dgpu_ensemble_synchronize(self.h), "dgpu_ensemble_synchronize")

    @property
    def points(self) -> int:
        return sum(c.points for c in self.cfgs)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            for m in self.members:
                m.h = None
            self.L.dgpu_ensemble_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

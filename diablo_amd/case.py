"""Channel case description: run-time decks, y-grid, wavenumbers and the
seeded initial state.

This is host-side set-up logic that sits in front of the time-stepper.  It
restates the parts of the reference's initialisation that the channel path
depends on:

* ``input.dat`` / ``input_chan.dat`` fixed-line decks
  (for/diablo_io.f:11-46, for/channel.f:1463-1509), read and written;
* the y-grid generator ``pre_process/create_grid_h5.m:28-75``;
* grid metrics from GY (for/hdf5.f:676-703, for/channel.f:1572-1652);
* INIT_FFT wavenumbers (for/fft.f:72-105) and INIT_CHAN_MPI index ranges at
  NPROCY = 1 (for/mpi.f:1080-1100);
* the RKW3 table (for/diablo_io.f:116-124);
* seeded initial conditions of the same forms as CREATE_FLOW_CHAN /
  CREATE_TH_CHAN (for/set_ics.f:2-259), with the compiler-specific
  RANDOM_NUMBER replaced by a counter-based generator (``uniform01``).

Arrays in "Fortran layout" are numpy arrays of shape (NY+2, NZ+2, NX+2),
C order, which is the reference's REAL*8 (0:NX+1, 0:NZ+1, 0:NY+1) column-major
array byte for byte; ``fourier_view`` gives the COMPLEX*16 (0:NX/2, 0:NZ+1,
0:NY+1) view of the same memory.
"""
from __future__ import annotations

import dataclasses
import math
import re
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

MAX_TH = 4


@dataclass
class ChannelConfig:
    """Everything the channel time-stepper reads from the decks and grid_def."""

    nx: int
    ny: int
    nz: int
    n_th: int = 1
    nu: float = 1.0e-3
    beta: float = 1.0
    lx: float = 2.0 * math.pi
    ly: float = 2.0
    lz: float = math.pi
    nu_v_scale: float = 1.0
    delta_t: float = 1.0e-3
    # VARIABLE_DT, CFL, UPDATE_DT (input.dat time-march line, for/diablo_io.f:33-34)
    variable_dt: bool = False
    cfl: float = 0.5
    update_dt: int = 1
    # driver settings of input.dat (for/diablo_io.f:31-36)
    create_new_flow: bool = True
    # USE_LES (input.dat line 2): the constant Smagorinsky model and the
    # eddy viscosity in the Crank-Nicolson terms (LES_MODEL_TYPE = 1)
    les: bool = False
    time_limit: float = 1.0e6
    reset_time: bool = False
    save_flow_int: int = 100000
    save_stats_int: int = 100000
    i_ro: float = 0.0
    grav_x: float = 0.0
    grav_y: float = 0.0
    grav_z: float = 0.0
    f_type: int = 3
    ubulk0: float = 0.0
    px0: float = 0.0
    omega0: float = 0.0
    amp_omega0: float = 0.0
    u_bc_ymin: int = 0
    u_bc_ymin_c1: float = 0.0
    v_bc_ymin: int = 0
    v_bc_ymin_c1: float = 0.0
    w_bc_ymin: int = 0
    w_bc_ymin_c1: float = 0.0
    u_bc_ymax: int = 0
    u_bc_ymax_c1: float = 0.0
    v_bc_ymax: int = 0
    v_bc_ymax_c1: float = 0.0
    w_bc_ymax: int = 0
    w_bc_ymax_c1: float = 0.0
    th_bc_ymin: List[int] = field(default_factory=lambda: [0])
    th_bc_ymin_c1: List[float] = field(default_factory=lambda: [0.0])
    th_bc_ymax: List[int] = field(default_factory=lambda: [0])
    th_bc_ymax_c1: List[float] = field(default_factory=lambda: [0.0])
    ri: List[float] = field(default_factory=lambda: [0.0])
    pr: List[float] = field(default_factory=lambda: [1.0])
    drhodx: List[float] = field(default_factory=lambda: [0.0])
    drhodz: List[float] = field(default_factory=lambda: [0.0])
    # y-grid (create_grid_h5.m inputs)
    grid_type: int = 3
    grid_cs: float = 1.75
    # initial condition
    ic: str = "channel"        # "channel": y(2-y) profile, "kh": tanh shear layer
    kick: float = 0.01
    streak: float = 0.0
    seed: int = 1234
    n_steps: int = 1

    # ---- derived sizes (for/header:21-28, for/header_mpi:13-14) ----
    @property
    def nxp(self) -> int:
        return self.nx // 2

    @property
    def nkx(self) -> int:
        return self.nx // 3

    @property
    def nkz(self) -> int:
        return self.nz // 3

    @property
    def tnkz(self) -> int:
        return 2 * (self.nz // 3)

    @property
    def points(self) -> int:
        return self.nx * self.ny * self.nz

    def field_shape(self):
        return (self.ny + 2, self.nz + 2, self.nx + 2)

    def replace(self, **kw) -> "ChannelConfig":
        return dataclasses.replace(self, **kw)

    def validate(self) -> None:
        for n in (self.nx, self.nz):
            if n < 4 or n % 2 or not fft_size_ok(n):
                raise ValueError(f"transform length {n} must be even and of the form 2^a 3^b 5^c")
        if self.ny < 4:
            raise ValueError("NY must be at least 4")
        if not 0 <= self.n_th <= MAX_TH:
            raise ValueError(f"N_TH must be in 0..{MAX_TH}")
        for name in ("th_bc_ymin", "th_bc_ymin_c1", "th_bc_ymax", "th_bc_ymax_c1", "ri", "pr", "drhodx", "drhodz"):
            if len(getattr(self, name)) < self.n_th:
                raise ValueError(f"{name} needs {self.n_th} entries")
        for code in (self.u_bc_ymin, self.w_bc_ymin, self.u_bc_ymax, self.w_bc_ymax):
            if code not in (0, 1):
                raise ValueError("U/W BC codes must be 0 or 1 (for/channel.f:2322)")
        for code in (self.v_bc_ymin, self.v_bc_ymax):
            if code not in (0, 1, 2):
                raise ValueError("V BC codes must be 0, 1 or 2 (for/channel.f:2383)")
        for n in range(self.n_th):
            if self.th_bc_ymin[n] not in (0, 1) or self.th_bc_ymax[n] not in (0, 1):
                raise ValueError("TH BC codes must be 0 or 1 (for/channel.f:2760)")


def fft_size_ok(n: int) -> bool:
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


# --------------------------------------------------------------------------
# Named workloads
# --------------------------------------------------------------------------

def channel_case(nx=64, ny=65, nz=64, re=180.0, dt=1.0e-3, kick=0.01, streak=0.0,
                 seed=1234, n_steps=200, n_th=1) -> ChannelConfig:
    """Closed channel, walls at y = 0 and 2 (tanh grid, CS = 1.75), laminar
    centreline velocity 1 held by a constant pressure gradient PX0 = -2 nu,
    passive scalar with Dirichlet 0 / 1 walls and Pr = 1."""
    nu = 1.0 / re
    return ChannelConfig(
        nx=nx, ny=ny, nz=nz, n_th=n_th, nu=nu, lx=2.0 * math.pi, ly=2.0, lz=math.pi,
        delta_t=dt, f_type=1, px0=-2.0 * nu,
        th_bc_ymin=[0] * n_th, th_bc_ymin_c1=[0.0] * n_th,
        th_bc_ymax=[0] * n_th, th_bc_ymax_c1=[1.0] * n_th,
        ri=[0.0] * n_th, pr=[1.0] * n_th, drhodx=[0.0] * n_th, drhodz=[0.0] * n_th,
        grid_type=3, grid_cs=1.75, ic="channel", kick=kick, streak=streak, seed=seed,
        n_steps=n_steps)


def kh_case(nx=256, ny=193, nz=256, re=1000.0, ri=0.1, dt=2.0e-3, kick=1.0e-3,
            seed=1234, n_steps=2000, lx=30.0, ly=20.0, lz=4.0) -> ChannelConfig:
    """Kelvin-Helmholtz shear layer U = theta = tanh(y) between free-slip walls
    (the KH_test deck: U/W Neumann, V Dirichlet, TH Neumann, GRAV_Y = 1)."""
    return ChannelConfig(
        nx=nx, ny=ny, nz=nz, n_th=1, nu=1.0 / re, lx=lx, ly=ly, lz=lz, delta_t=dt,
        grav_y=1.0, f_type=3,
        u_bc_ymin=1, v_bc_ymin=0, w_bc_ymin=1, u_bc_ymax=1, v_bc_ymax=0, w_bc_ymax=1,
        th_bc_ymin=[1], th_bc_ymin_c1=[0.0], th_bc_ymax=[1], th_bc_ymax_c1=[0.0],
        ri=[ri], pr=[1.0], drhodx=[0.0], drhodz=[0.0],
        grid_type=2, grid_cs=1.5, ic="kh", kick=kick, seed=seed, n_steps=n_steps)


def ensemble_cases(n: int = 64, seed: int = 2024, dt: float = 2.0e-3, kick: float = 1.0e-3):
    """Kelvin-Helmholtz members on uneven grids: NX = NZ in {64, 96, 128, 192}
    and NY in {33, 65, 97, 129} (16 shapes, used in turn), Ri uniform in
    [0.05, 0.25], Re in {500, 1000} and member seed = seed + m."""
    rng = np.random.default_rng(seed)
    shapes = [(nx, ny) for nx in (64, 96, 128, 192) for ny in (33, 65, 97, 129)]
    out = []
    for m in range(n):
        nx, ny = shapes[m % len(shapes)]
        ri = float(rng.uniform(0.05, 0.25))
        re = float(rng.choice([500.0, 1000.0]))
        out.append(kh_case(nx, ny, nx, re=re, ri=ri, dt=dt, kick=kick, seed=seed + m, n_steps=500))
    return out


WORKLOADS = {
    # CPU-runnable channel, FP64, 200 steps
    "channel64": lambda: channel_case(64, 65, 64, re=180.0, dt=1.0e-3, kick=0.01, n_steps=200),
    # the large channel: 384 x 257 x 384, Re = 395, streak seed, 1000 steps
    "channel384": lambda: channel_case(384, 257, 384, re=395.0, dt=5.0e-4, kick=0.05, streak=0.1,
                                       n_steps=1000),
    # Kelvin-Helmholtz shear layer 256 x 193 x 256
    "kh256": lambda: kh_case(256, 193, 256, re=1000.0, ri=0.1, dt=2.0e-3, kick=1.0e-3, n_steps=2000),
}


# --------------------------------------------------------------------------
# y-grid: pre_process/create_grid_h5.m:28-75
# --------------------------------------------------------------------------

def create_grid_gy(grid_type: int, n: int, length: float, cs: float) -> np.ndarray:
    """GY(1..N+1) as written to /grids/y by create_grid_h5.m."""
    G = np.zeros(n + 3)
    GF = np.zeros(n + 3)
    for J in range(1, n + 2):
        x = (2.0 * (J - 1)) / n - 1.0
        if grid_type == 1:
            G[J + 1] = (length / 2.0) * math.tanh(cs * x) / math.tanh(cs)
        elif grid_type == 2:
            G[J + 1] = (length / 2.0) * (cs * x ** 3 + x) / (cs + 1)
        elif grid_type == 3:
            G[J + 1] = (length / 2.0) * math.tanh(cs * x) / math.tanh(cs) + length / 2.0
        else:
            raise ValueError("grid type must be 1, 2 or 3")
    for J in range(1, n + 1):
        GF[J + 1] = (G[J + 1] + G[J + 2]) / 2.0
    gf_lower, gf_upper = GF[2], GF[n + 1]
    g_lower, g_upper = G[2], G[n + 2]
    for J in range(1, n + 1):
        GF[J + 1] = GF[J + 1] * (g_upper - g_lower) / (gf_upper - gf_lower)
    for J in range(1, n + 2):
        G[J + 1] = G[J + 1] * (g_upper - g_lower) / (gf_upper - gf_lower)
    gf_lower = GF[2]
    shift = g_lower - gf_lower
    for J in range(1, n + 1):
        GF[J + 1] = GF[J + 1] + shift
    for J in range(1, n + 2):
        G[J + 1] = G[J + 1] + shift
    return G[2:n + 3].copy()


def grid_gy(cfg: ChannelConfig) -> np.ndarray:
    return create_grid_gy(cfg.grid_type, cfg.ny, cfg.ly, cfg.grid_cs)


def grid_metrics(gy_1: np.ndarray, ny: int):
    """GY, GYF, DY, DYF over 0:NY+1 from GY(1..NY+1).

    ReadGridHDF5 (for/hdf5.f:676-703) and CREATE_GRID_CHAN
    (for/channel.f:1646-1652); GY(0), DY(0) and DYF(0) are never set (0)."""
    GY = np.zeros(ny + 2)
    GYF = np.zeros(ny + 2)
    DY = np.zeros(ny + 2)
    DYF = np.zeros(ny + 2)
    GY[1:ny + 2] = gy_1[:ny + 1]
    for j in range(1, ny + 1):
        GYF[j] = 0.5 * (GY[j] + GY[j + 1])
    GYF[0] = 2.0 * GYF[1] - GYF[2]
    GYF[ny + 1] = 2.0 * GYF[ny] - GYF[ny - 1]
    for j in range(1, ny + 2):
        DY[j] = GYF[j] - GYF[j - 1]
    for j in range(1, ny + 1):
        DYF[j] = GY[j + 1] - GY[j]
    DYF[ny + 1] = DYF[ny]
    return GY, GYF, DY, DYF


def wavenumbers(cfg: ChannelConfig):
    """KX(0:NXP), KZ(0:TNKZ), KX2, KZ2 exactly as INIT_FFT (for/fft.f:72-105).

    PI = 4*atan(1d0); KX(I) = dble(I)*(2*PI)/LX for I < NXP (KX(NXP) = 0);
    KZ(K) = K*(2*PI)/LZ for K <= NKZ, KZ(TNKZ+1-K) = -K*(2*PI)/LZ."""
    pi = 4.0 * math.atan(1.0)
    kx = np.zeros(cfg.nxp + 1)
    for i in range(cfg.nxp):
        kx[i] = (float(i) * (2.0 * pi)) / cfg.lx
    kz = np.zeros(cfg.tnkz + 1)
    for k in range(cfg.nkz + 1):
        kz[k] = (float(k) * (2.0 * pi)) / cfg.lz
    for k in range(1, cfg.nkz + 1):
        kz[cfg.tnkz + 1 - k] = (float(-k) * (2.0 * pi)) / cfg.lz
    return kx, kz, kx * kx, kz * kz


def kz_of_slot(cfg: ChannelConfig) -> np.ndarray:
    """Integer z-wavenumber held by packed slot K = 0..TNKZ (for/mpi.f:1602-1606)."""
    s = np.arange(cfg.tnkz + 1)
    return np.where(s <= cfg.nkz, s, s - (cfg.tnkz + 1))


def index_ranges(cfg: ChannelConfig):
    """JSTART, JEND, JSTART_TH, JEND_TH at NPROCY = 1 (for/mpi.f:1080-1100)."""
    jstart_th = [1 if cfg.th_bc_ymin[n] == 1 else 2 for n in range(cfg.n_th)]
    return 2, cfg.ny, jstart_th, [cfg.ny] * cfg.n_th


def rkw3_table(dt: float):
    """H_BAR, BETA_BAR, ZETA_BAR (for/diablo_io.f:116-124)."""
    h = np.array([dt * (8.0 / 15.0), dt * (2.0 / 15.0), dt * (5.0 / 15.0)])
    b = np.array([1.0, 25.0 / 8.0, 9.0 / 4.0])
    z = np.array([0.0, -17.0 / 8.0, -5.0 / 4.0])
    return h, b, z


# --------------------------------------------------------------------------
# Seeded random numbers and initial conditions
# --------------------------------------------------------------------------

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, component: int, shape) -> np.ndarray:
    """Counter-based uniform [0,1) numbers, one per (component, j, k, i).

    xi = (splitmix64(seed * 0x9E3779B97F4A7C15 + counter) >> 11) * 2^-53 with
    counter = ((component * NJ + j) * NK + k) * NI + i.  The result does not
    depend on thread count, ordering or platform, which the reference's
    RANDOM_NUMBER (for/set_ics.f:12-17) does not guarantee."""
    nj, nk, ni = shape
    cnt = np.arange(nj * nk * ni, dtype=np.uint64) + np.uint64(component * nj * nk * ni)
    with np.errstate(over="ignore"):
        key = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * np.uint64(0x9E3779B97F4A7C15) + cnt
    bits = _splitmix64(key) >> np.uint64(11)
    return (bits.astype(np.float64) * (1.0 / 9007199254740992.0)).reshape(shape)


def initial_physical(cfg: ChannelConfig, gy_1: Optional[np.ndarray] = None):
    """Physical-space start fields in Fortran layout: (U1, U2, U3, [TH_n]).

    channel: U1 = y(2-y) + kick*(xi-0.5)*y(2-y) [+ streak*sin(4 pi z/LZ)*y(2-y)],
             V, W = kick*(xi-0.5)*y(2-y), on J = 0..NY+1 (for/set_ics.f:37-93
             with the noise made to vanish at the walls);
             TH linear between the wall values on J = 1..NY.
    kh:      U1 = tanh(GYF) on J = 0..NY (for/set_ics.f:63-73) plus
             kick*(xi-0.5) on every component and J = 0..NY+1
             (for/set_ics.f:82-93); TH = tanh(GYF) on J = 1..NY
             (for/set_ics.f:237-245)."""
    if gy_1 is None:
        gy_1 = grid_gy(cfg)
    _, GYF, _, _ = grid_metrics(gy_1, cfg.ny)
    shp = cfg.field_shape()
    NX, NY, NZ = cfg.nx, cfg.ny, cfg.nz
    u = [np.zeros(shp) for _ in range(3)]
    ths = [np.zeros(shp) for _ in range(cfg.n_th)]
    noise_shape = (NY + 2, NZ, NX)
    xi = [uniform01(cfg.seed, c, noise_shape) - 0.5 for c in range(3)]
    y = GYF[:, None, None]
    if cfg.ic == "channel":
        env = y * (2.0 - y)
        z = (np.arange(NZ) * cfg.lz / NZ)[None, :, None]
        base = np.broadcast_to(env, noise_shape).copy()
        u[0][:, :NZ, :NX] = base + cfg.kick * xi[0] * env + cfg.streak * np.sin(4.0 * math.pi * z / cfg.lz) * env
        u[1][:, :NZ, :NX] = cfg.kick * xi[1] * env
        u[2][:, :NZ, :NX] = cfg.kick * xi[2] * env
        for n in range(cfg.n_th):
            lo, hi = cfg.th_bc_ymin_c1[n], cfg.th_bc_ymax_c1[n]
            prof = lo + (hi - lo) * GYF[1:NY + 1] / cfg.ly
            ths[n][1:NY + 1, :NZ, :NX] = prof[:, None, None]
    elif cfg.ic == "kh":
        t = np.tanh(GYF)
        u[0][0:NY + 1, :NZ, :NX] = t[0:NY + 1, None, None]
        for c in range(3):
            u[c][:, :NZ, :NX] += cfg.kick * xi[c]
        for n in range(cfg.n_th):
            ths[n][1:NY + 1, :NZ, :NX] = t[1:NY + 1, None, None]
    else:
        raise ValueError(f"unknown initial condition {cfg.ic!r}")
    return u[0], u[1], u[2], ths


def fourier_view(a: np.ndarray) -> np.ndarray:
    """COMPLEX*16 (NY+2, NZ+2, NX/2+1) view of a Fortran-layout REAL*8 field."""
    return a.view(np.complex128)


def kept(cfg: ChannelConfig, a: np.ndarray) -> np.ndarray:
    """The kept Fourier slots I = 0..NKX, K = 0..TNKZ of every plane."""
    return fourier_view(a)[:, :cfg.tnkz + 1, :cfg.nkx + 1]


# --------------------------------------------------------------------------
# Decks: input.dat (for/diablo_io.f:17-46) and input_chan.dat
# (for/channel.f:1466-1509)
# --------------------------------------------------------------------------

def _fl(x) -> str:
    return repr(float(x))


def _lg(b) -> str:
    return ".TRUE." if b else ".FALSE."


def write_decks(cfg: ChannelConfig, directory: str, save_flow_int: Optional[int] = None,
                save_stats_int: Optional[int] = None) -> None:
    """Writes input.dat and input_chan.dat in the reference's fixed-line layout
    (and grid_def with the sizes)."""
    save_flow_int = cfg.save_flow_int if save_flow_int is None else save_flow_int
    save_stats_int = cfg.save_stats_int if save_stats_int is None else save_stats_int
    lines = [
        "C A data file for diablo.  This is a generic, easy method of data",
        "C input, as long as you don't change the number of comment lines.",
        "C Note that the grid size is defined in the file grid_def.",
        "C   FLAVOR        VERSION",
        "    'Basic'        2.0",
        "C   USE_MPI       USE_LES",
        f"    .TRUE.       {_lg(cfg.les)}",
        "C Parameters:     NU,        BETA,      LX,      LY,     LZ",
        f"   {_fl(cfg.nu)} {_fl(cfg.beta)} {_fl(cfg.lx)} {_fl(cfg.ly)} {_fl(cfg.lz)}",
        "C Vertical viscosity scale factor NU_V_SCALE (NU_V=NU*NU_V_SCALE)",
        f"   {_fl(cfg.nu_v_scale)}",
        "C Method:     NUM_PER_DIR, CREATE_NEW_FLOW",
        f"   2   {_lg(cfg.create_new_flow)}",
        "C Time March:  N_TIME_STEPS, TIME_LIMIT, DELTA_T, RESET_TIME, VARIABLE_DT, CFL, UPDATE_DT",
        f"   {cfg.n_steps}  {_fl(cfg.time_limit)}  {_fl(cfg.delta_t)}  {_lg(cfg.reset_time)}  {_lg(cfg.variable_dt)}"
        f"  {_fl(cfg.cfl)}  {cfg.update_dt}",
        "C I/O:         VERBOSITY, SAVE_FLOW_INT, SAVE_STATS_INT,  MOVIE",
        f"   1  {save_flow_int}  {save_stats_int}  .FALSE.",
        "C Here include 6*N_TH lines, see below for format",
    ]
    for n in range(cfg.n_th):
        lines += [
            "C         CREATE_FLOW_TH(1)  Create new field or Read from DIABLO_TH.START",
            "   .TRUE.",
            "C         FILTER_TH(1)         FILTER_INT(1)     (If and how often to filter)",
            "   .FALSE.   10",
            "C         RI_TAU(1)          PR(1)",
            f"   {_fl(cfg.ri[n])}  {_fl(cfg.pr[n])}",
        ]
    with open(os.path.join(directory, "input.dat"), "w") as f:
        f.write("\n".join(lines) + "\n")
    ic_type = 3 if cfg.ic == "kh" else 1
    lines = [
        "C A data file for diablo.  This is a generic, easy method of data",
        "C input, as long as you don't change the number of comment lines.",
        "C Note that the grid size is defined in the file grid_def.",
        "C VERSION",
        "    2.0",
        "C TIME_AD_METH",
        "    1",
        "C LES:   LES_MODEL_TYPE",
        "    1",
        "C ICs:    IC_TYPE,    KICK",
        f"    {ic_type}  {_fl(cfg.kick)}",
        "C Rotation:  I_RO (or Coriolis parameter, f)",
        f"    {_fl(cfg.i_ro)}",
        "C Upward vertical vector components, GRAV_X, GRAV_Y, GRAV_Z",
        f"    {_fl(cfg.grav_x)}  {_fl(cfg.grav_y)}  {_fl(cfg.grav_z)}",
        "C Forcing:      F_TYPE,    UBULK0,   PX0,    OMEGA0,   AMP_OMEGA0",
        f"    {cfg.f_type}  {_fl(cfg.ubulk0)}  {_fl(cfg.px0)}  {_fl(cfg.omega0)}  {_fl(cfg.amp_omega0)}",
    ]
    for name in ("u", "v", "w"):
        lines += [f"C BCs: {name.upper()}_BC_YMIN, {name.upper()}_BC_YMIN_C1, C2, C3",
                  f"    {getattr(cfg, name + '_bc_ymin')}  {_fl(getattr(cfg, name + '_bc_ymin_c1'))}  0.0  0.0"]
    for name in ("u", "v", "w"):
        lines += [f"C BCs: {name.upper()}_BC_YMAX, {name.upper()}_BC_YMAX_C1, C2, C3",
                  f"    {getattr(cfg, name + '_bc_ymax')}  {_fl(getattr(cfg, name + '_bc_ymax_c1'))}  0.0  0.0"]
    lines.append("C Here include 6*N_TH lines, see below for format")
    for n in range(cfg.n_th):
        lines += [
            f"C BCs: TH_BC_YMIN({n + 1}), TH_BC_YMIN_C1, C2, C3",
            f"    {cfg.th_bc_ymin[n]}  {_fl(cfg.th_bc_ymin_c1[n])}  0.0  0.0",
            f"C BCs: TH_BC_YMAX({n + 1}), TH_BC_YMAX_C1, C2, C3",
            f"    {cfg.th_bc_ymax[n]}  {_fl(cfg.th_bc_ymax_c1[n])}  0.0  0.0",
            "C Background scalar gradients, DRHODX, DRHODZ",
            f"    {_fl(cfg.drhodx[n])}  {_fl(cfg.drhodz[n])}",
        ]
    with open(os.path.join(directory, "input_chan.dat"), "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(os.path.join(directory, "grid_def"), "w") as f:
        f.write("!----*|--.---------.---------.---------.---------.---------.---------.-|-------|\n"
                "! grid_def, the portion of the header that defines the grid size for diablo.\n"
                "!----*|--.---------.---------.---------.---------.---------.---------.-|-------|\n"
                f"      PARAMETER (NX={cfg.nx})\n      PARAMETER (NY={cfg.ny})\n"
                f"      PARAMETER (NZ={cfg.nz})\n      PARAMETER (N_TH={cfg.n_th})\n")


def read_grid_def(directory: str):
    """NX, NY, NZ, N_TH from grid_def's PARAMETER statements."""
    vals = {}
    with open(os.path.join(directory, "grid_def")) as f:
        for line in f:
            if line.lstrip().startswith("!"):
                continue
            m = re.search(r"PARAMETER\s*\(\s*(\w+)\s*=\s*(\d+)\s*\)", line, re.IGNORECASE)
            if m:
                vals[m.group(1).upper()] = int(m.group(2))
    return vals["NX"], vals["NY"], vals["NZ"], vals.get("N_TH", 0)


class _ListDirected:
    """Minimal Fortran list-directed reader: READ(u,*) with no items skips one
    record; with items it takes values from the next record(s)."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.pos = 0

    def skip(self):
        self.pos += 1

    def read(self, n: int):
        vals: list = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise ValueError("end of deck")
            line = self.lines[self.pos]
            self.pos += 1
            toks = line.replace(",", " ").split()
            vals.extend(toks[: n - len(vals)])
        return vals


def _logical(tok: str) -> bool:
    t = tok.strip().upper().lstrip(".")
    return t.startswith("T")


def read_decks(directory: str, nx: int, ny: int, nz: int, n_th: int) -> ChannelConfig:
    """Reads input.dat and input_chan.dat (sizes come from grid_def)."""
    with open(os.path.join(directory, "input.dat")) as f:
        r = _ListDirected(f.read())
    for _ in range(4):
        r.skip()
    _, version = r.read(2)
    if float(version) != 2.0:
        raise ValueError("Wrong input data format.")
    r.skip()
    _use_mpi, les = r.read(2)
    r.skip()
    nu, beta, lx, ly, lz = map(float, r.read(5))
    r.skip()
    (nu_v_scale,) = map(float, r.read(1))
    r.skip()
    num_per_dir, create = r.read(2)
    if int(num_per_dir) != 2:
        raise ValueError("only NUM_PER_DIR = 2 (channel) is supported")
    r.skip()
    nsteps, tlim, dt, reset, vdt, cfl, upd = r.read(7)
    r.skip()
    _verbosity, sfi, ssi, _movie = r.read(4)
    r.skip()
    ri, pr = [], []
    for _ in range(n_th):
        r.skip(); r.read(1); r.skip(); r.read(2); r.skip()
        a, b = map(float, r.read(2))
        ri.append(a)
        pr.append(b)
    with open(os.path.join(directory, "input_chan.dat")) as f:
        c = _ListDirected(f.read())
    for _ in range(4):
        c.skip()
    (cv,) = c.read(1)
    if float(cv) != 2.0:
        raise ValueError("Wrong input data format input_chan")
    c.skip()
    (tam,) = c.read(1)
    if int(tam) != 1:
        raise ValueError("only TIME_AD_METH = 1 (RK_CHAN_1) is supported")
    c.skip()
    (les_model_type,) = c.read(1)
    if _logical(les) and int(les_model_type) != 1:
        # types 2 and 3 stop in the reference's MPI build (for/les.f:272-276)
        raise ValueError("only LES_MODEL_TYPE = 1 (constant Smagorinsky) is supported")
    c.skip()
    ic_type, kick = c.read(2)
    c.skip()
    (i_ro,) = map(float, c.read(1))
    c.skip()
    gx, gyv, gz = map(float, c.read(3))
    c.skip()
    f_type, ubulk0, px0, omega0, amp = c.read(5)
    bc = {}
    for key in ("u_bc_ymin", "v_bc_ymin", "w_bc_ymin", "u_bc_ymax", "v_bc_ymax", "w_bc_ymax"):
        c.skip()
        code, c1, _c2, _c3 = c.read(4)
        bc[key] = int(code)
        bc[key + "_c1"] = float(c1)
    c.skip()
    thmin, thmin1, thmax, thmax1, drx, drz = [], [], [], [], [], []
    for _ in range(n_th):
        c.skip()
        a, b, _, _ = c.read(4)
        thmin.append(int(a)); thmin1.append(float(b))
        c.skip()
        a, b, _, _ = c.read(4)
        thmax.append(int(a)); thmax1.append(float(b))
        c.skip()
        a, b = map(float, c.read(2))
        drx.append(a); drz.append(b)
    return ChannelConfig(
        nx=nx, ny=ny, nz=nz, n_th=n_th, nu=nu, beta=beta, lx=lx, ly=ly, lz=lz,
        nu_v_scale=nu_v_scale, delta_t=float(dt), variable_dt=_logical(vdt), cfl=float(cfl),
        update_dt=int(upd), create_new_flow=_logical(create), les=_logical(les), time_limit=float(tlim),
        reset_time=_logical(reset),
        save_flow_int=int(sfi), save_stats_int=int(ssi), i_ro=i_ro, grav_x=gx, grav_y=gyv, grav_z=gz,
        f_type=int(f_type), ubulk0=float(ubulk0), px0=float(px0), omega0=float(omega0),
        amp_omega0=float(amp), th_bc_ymin=thmin, th_bc_ymin_c1=thmin1, th_bc_ymax=thmax,
        th_bc_ymax_c1=thmax1, ri=ri, pr=pr, drhodx=drx, drhodz=drz,
        ic="kh" if int(ic_type) == 3 else "channel", kick=float(kick), n_steps=int(nsteps), **bc)

"""Pins of RK_CHAN_1's explicit terms (for/channel.f:123-290, 410-720) and
of the Crank-Nicolson halves and the pressure update around them
(for/channel.f:606-626, 740-746, 1059-1065), for the oracle's restatement.

Each pin drives one substep (or step) of an oracle built by ``make(cfg,
gy_1=None)`` from a state chosen so that the result has a closed form that
does not come from re-evaluating the oracle's own expression: single Fourier
modes whose self-products fall outside the kept band, linear profiles for
which the staggered averages are exact, a y-independent flow evaluated with
numpy's FFT, conservation by telescoping flux differences, discrete
eigenvectors and the exact discrete projection of a pressure gradient.

RK_CHAN_1 leaves the explicit terms of the substep in CF1-3 and CFTH
(for/channel.f:170-290, 410-571, 665-720: nothing after them writes CF), so
after RK_STEP = 1 (ZETA_BAR(1) = 0, no CF of an earlier stage) CF holds A(U)
of the substep's input U.

tests/test_oracle_rk_terms.py runs every pin on the oracle and on copies of
it built with one planted mistake each (PLANT sites in diablo_oracle.c),
and requires each copy to be rejected.
"""
import math

import numpy as np

from diablo_amd import case

U1, U2, U3, P, F1, F2, F3 = 0, 1, 2, 3, 7, 8, 9
CTH0, CFTH = 40, 48
GY, GYF, DY, DYF = 0, 1, 2, 3

RK_H = np.array([8.0, 2.0, 5.0]) / 15.0          # H_BAR / DELTA_T (for/diablo_io.f:116-124)
RK_BETA = np.array([1.0, 25.0 / 8.0, 9.0 / 4.0])
RK_ZETA = np.array([0.0, -17.0 / 8.0, -5.0 / 4.0])


class PinFailure(AssertionError):
    pass


def check(ok, msg):
    if not ok:
        raise PinFailure(msg)


def fv(o, fid):
    return case.fourier_view(o.field(fid))


def kx_of(cfg, i):
    return i * (2.0 * math.pi / cfg.lx)


def kz_of(cfg, k):
    s = k if k <= cfg.nkz else k - (cfg.tnkz + 1)
    return s * (2.0 * math.pi / cfg.lz)


def conj_slot(cfg, k):
    return 0 if k == 0 else cfg.tnkz + 1 - k


def set_mode(cfg, arr, i, k, prof):
    """Mode (i, k) of a real field: for i = 0, k != 0 the -kz slot holds the
    conjugate (for/mpi.f:1602-1606 packing)."""
    arr[:, k, i] = prof
    if i == 0 and k != 0:
        arr[:, conj_slot(cfg, k), 0] = np.conj(prof)


def crand(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def rel(got, want):
    scale = max(float(np.abs(want).max()), 1e-300)
    return float(np.abs(np.asarray(got) - np.asarray(want)).max()) / scale


def expect(name, got, want, tol):
    e = rel(got, want)
    check(e <= tol, f"{name}: rel err {e:.3e} > {tol:.0e}")
    return e


# ---------------------------------------------------------------------------
# explicit linear terms: Coriolis, horizontal viscosity and diffusivity with
# BETA, GRAV_X / GRAV_Z buoyancy, background scalar gradients
# ---------------------------------------------------------------------------
def pin_linear_terms(make):
    """One velocity/scalar mode whose self-products fall outside the kept band
    (kx index 5 at NX = 24, NKX = 8: products at 0 and 10; kz index 4 at
    NZ = 18, NKZ = 6: products at 0 and +-8; the mean part is
    differentiated by i kx = i kz = 0), so CF holds only the linear terms of
    for/channel.f:175-200, 229-243, 273-290:
      CF1 = -NU k^(2 BETA) U1 + I_RO U3 + RI GRAV_X TH
      CF3 = -NU k^(2 BETA) U3 - I_RO U1 + RI GRAV_Z TH
      CFTH = -(NU/PR) k^(2 BETA) TH - DRHODX U1 - DRHODZ U3,
    on rows JSTART..JEND (2..NY), CF2 = 0, every other slot zero."""
    worst = 0.0
    for beta in (1.0, 2.0):
        cfg = case.channel_case(24, 17, 18, kick=0.0).replace(
            i_ro=0.37, beta=beta, ri=[0.6], grav_x=0.3, grav_y=0.0, grav_z=-0.45,
            drhodx=[0.21], drhodz=[-0.13], pr=[0.7], nu=0.011)
        for (i, k) in ((5, 0), (0, 4)):
            o = make(cfg)
            rng = np.random.default_rng(100 + i + k + int(beta))
            ny2 = cfg.ny + 2
            a, c, t = crand(rng, ny2), crand(rng, ny2), crand(rng, ny2)
            set_mode(cfg, fv(o, U1), i, k, a)
            set_mode(cfg, fv(o, U3), i, k, c)
            set_mode(cfg, fv(o, CTH0), i, k, t)
            o.rk_chan_1(1)
            k2 = kx_of(cfg, i) ** 2 + kz_of(cfg, k) ** 2
            lv, lt = cfg.nu * k2 ** beta, cfg.nu / cfg.pr[0] * k2 ** beta
            rows = slice(2, cfg.ny + 1)
            want = {F1: -lv * a + cfg.i_ro * c + cfg.ri[0] * cfg.grav_x * t,
                    F3: -lv * c - cfg.i_ro * a + cfg.ri[0] * cfg.grav_z * t,
                    CFTH: -lt * t - cfg.drhodx[0] * a - cfg.drhodz[0] * c}
            for fid, w in want.items():
                full = np.zeros_like(fv(o, fid)[:, :cfg.tnkz + 1, :cfg.nxp])
                full[rows, k, i] = w[rows]
                if i == 0 and k != 0:
                    full[rows, conj_slot(cfg, k), 0] = np.conj(w[rows])
                got = fv(o, fid)[:, :cfg.tnkz + 1, :cfg.nxp]
                worst = max(worst, expect(f"linear terms beta={beta} mode=({i},{k}) field {fid}", got, full, 1e-13))
            cf2 = fv(o, F2)[:, :cfg.tnkz + 1, :cfg.nxp]
            check(np.abs(cf2).max() <= 1e-13 * np.abs(a).max(), "CF2 must stay zero")
    return worst


# ---------------------------------------------------------------------------
# buoyancy on GY (for/channel.f:209-226)
# ---------------------------------------------------------------------------
def pin_buoyancy(make):
    """TH = s GYF + r in a non-mean mode: the DYF-weighted average of
    for/channel.f:216-217 (TH(J) weighted by DYF(J-1)) is exact linear
    interpolation, so CF2 = RI GRAV_Y (s GY(J) + r) on J = 2..NY on the
    stretched grid; the (0,0) mode of TH (the plane mean balanced by the
    hydrostatic pressure) adds nothing; rows 0, 1, NY+1 stay zero."""
    cfg = case.channel_case(24, 33, 18, kick=0.0).replace(ri=[0.6], grav_y=1.3)
    o = make(cfg)
    gy, gyf = o.vec(GY), o.vec(GYF)
    rng = np.random.default_rng(7)
    (s1, r1, s2, r2) = crand(rng, 4)
    th = fv(o, CTH0)
    th[:, 0, 0] = 3.0 * gyf - 0.4
    set_mode(cfg, th, 1, 0, s1 * gyf + r1)
    set_mode(cfg, th, 0, 2, s2 * gyf + r2)
    o.rk_chan_1(1)
    g = cfg.ri[0] * cfg.grav_y
    full = np.zeros((cfg.ny + 2, cfg.tnkz + 1, cfg.nxp), complex)
    rows = slice(2, cfg.ny + 1)
    full[rows, 0, 1] = g * (s1 * gy[rows] + r1)
    full[rows, 2, 0] = g * (s2 * gy[rows] + r2)
    full[rows, conj_slot(cfg, 2), 0] = np.conj(full[rows, 2, 0])
    got = fv(o, F2)[:, :cfg.tnkz + 1, :cfg.nxp]
    return expect("buoyancy on GY", got, full, 1e-13)


# ---------------------------------------------------------------------------
# products with U2 (for/channel.f:468-571, 703-720), stretched grid
# ---------------------------------------------------------------------------
def pin_v_products_stretched(make):
    """U2 = c GY + v0 (mean mode), U1 = alpha + (p GYF + q) e^{i kx x},
    U3 = gamma + (p3 GYF + q3) e^{i kz z}, TH = beta0, on the tanh grid.
      * U1 on GY weighted by DYF (for/channel.f:476-477): for a linear profile
        the result is the mean over the two cells around GY(J),
        p (GY(J-1) + GY(J+1))/2 + q, so CF2(1,0) = -i kx V(J) (that), and the
        same for U3 in CF2(0,1) (for/channel.f:489-498);
      * d/dy(V^2) (for/channel.f:552-571): (V(J)+V(J+1))/2 = c GYF(J) + v0,
        so CF2(0,0) = -c (c (GYF(J) + GYF(J-1)) + 2 v0);
      * flux forms of a constant times V (for/channel.f:510-550, 703-720):
        CF1(0,0) = -alpha c, CF3(0,0) = -gamma c, CFTH(0,0) = -beta0 c;
      * U1 TH and U3 TH (for/channel.f:670-697): CFTH(1,0) = -i kx beta0 a,
        CFTH(0,1) = -i kz beta0 b."""
    cfg = case.channel_case(24, 33, 18, kick=0.0)
    o = make(cfg)
    gy, gyf = o.vec(GY), o.vec(GYF)
    rng = np.random.default_rng(11)
    p, q, p3, q3 = crand(rng, 4)
    c, v0, alpha, gamma, beta0 = 0.7, -0.3, 0.45, -0.6, 1.7
    ny = cfg.ny
    v = c * gy + v0
    v[0] = v0
    fv(o, U2)[:, 0, 0] = v
    a, b = p * gyf + q, p3 * gyf + q3
    fv(o, U1)[:, 0, 0] = alpha
    set_mode(cfg, fv(o, U1), 1, 0, a)
    fv(o, U3)[:, 0, 0] = gamma
    set_mode(cfg, fv(o, U3), 0, 1, b)
    fv(o, CTH0)[:, 0, 0] = beta0
    o.rk_chan_1(1)
    kx, kz = kx_of(cfg, 1), kz_of(cfg, 1)
    J = np.arange(2, ny + 1)
    mid = 0.5 * (gy[J - 1] + gy[J + 1])
    worst = 0.0
    worst = max(worst, expect("U1 on GY x V", fv(o, F2)[J, 0, 1], -1j * kx * v[J] * (p * mid + q), 1e-13))
    worst = max(worst, expect("U3 on GY x V", fv(o, F2)[J, 1, 0], -1j * kz * v[J] * (p3 * mid + q3), 1e-13))
    worst = max(worst, expect("d/dy V^2", fv(o, F2)[J, 0, 0], -c * (c * (gyf[J] + gyf[J - 1]) + 2 * v0), 1e-13))
    worst = max(worst, expect("d/dy(U1 V)", fv(o, F1)[J, 0, 0], np.full(J.size, -alpha * c), 1e-13))
    worst = max(worst, expect("d/dy(U3 V)", fv(o, F3)[J, 0, 0], np.full(J.size, -gamma * c), 1e-13))
    worst = max(worst, expect("d/dy(TH V)", fv(o, CFTH)[J, 0, 0], np.full(J.size, -beta0 * c), 1e-13))
    worst = max(worst, expect("d/dx(U1 TH)", fv(o, CFTH)[J, 0, 1], -1j * kx * beta0 * a[J], 1e-13))
    worst = max(worst, expect("d/dz(U3 TH)", fv(o, CFTH)[J, 1, 0], -1j * kz * beta0 * b[J], 1e-13))
    for fid in (F1, F2, F3, CFTH):
        arr = fv(o, fid)[:, :cfg.tnkz + 1, :cfg.nxp]
        check(np.abs(arr[[0, 1, ny + 1]]).max() == 0.0, f"rows 0, 1, NY+1 of field {fid} must stay zero")
    return worst


def pin_v_products_uniform(make):
    """On a uniform grid the simple averages of the flux forms are exact for
    linear profiles: with U1 = alpha + a(y) e^{i kx x}, a = p GYF + q,
    U2 = V = c GY + v0, U3 = gamma + b(y) e^{i kz z}, TH = beta0 + t(y) e^{i kx x},
    the face flux V (a(J) + a(J-1))/2 is g(GY(J)) for the quadratic
    g = V a, and its difference over DYF is g'(GYF(J)) = 2 c p GYF + c q + p v0.
      CF1(1,0) = -NU kx^2 a - i kx (2 alpha a) - g'_a
      CF3(0,1) = -NU kz^2 b - i kz (2 gamma b) - g'_b
      CFTH(1,0) = -(NU/PR) kx^2 t - i kx (alpha t + beta0 a) - g'_t
    (for/channel.f:175-189, 410-465, 510-550, 670-720)."""
    cfg = case.channel_case(24, 33, 18, kick=0.0).replace(pr=[0.8])
    gy_1 = np.linspace(0.0, 2.0, cfg.ny + 1)
    o = make(cfg, gy_1)
    gy, gyf = o.vec(GY), o.vec(GYF)
    rng = np.random.default_rng(12)
    p, q, p3, q3, pt, qt = crand(rng, 6)
    c, v0, alpha, gamma, beta0 = -0.8, 0.25, 0.35, 0.5, -1.1
    ny = cfg.ny
    v = c * gy + v0
    v[0] = v0
    fv(o, U2)[:, 0, 0] = v
    a, b, t = p * gyf + q, p3 * gyf + q3, pt * gyf + qt
    fv(o, U1)[:, 0, 0] = alpha
    set_mode(cfg, fv(o, U1), 1, 0, a)
    fv(o, U3)[:, 0, 0] = gamma
    set_mode(cfg, fv(o, U3), 0, 1, b)
    fv(o, CTH0)[:, 0, 0] = beta0
    set_mode(cfg, fv(o, CTH0), 1, 0, t)
    o.rk_chan_1(1)
    kx, kz = kx_of(cfg, 1), kz_of(cfg, 1)
    J = np.arange(2, ny + 1)

    def gp(pp, qq):
        return 2 * c * pp * gyf[J] + c * qq + pp * v0

    worst = 0.0
    worst = max(worst, expect("CF1(1,0) uniform", fv(o, F1)[J, 0, 1],
                              -cfg.nu * kx ** 2 * a[J] - 1j * kx * 2 * alpha * a[J] - gp(p, q), 1e-12))
    worst = max(worst, expect("CF3(0,1) uniform", fv(o, F3)[J, 1, 0],
                              -cfg.nu * kz ** 2 * b[J] - 1j * kz * 2 * gamma * b[J] - gp(p3, q3), 1e-12))
    worst = max(worst, expect("CFTH(1,0) uniform", fv(o, CFTH)[J, 0, 1],
                              -cfg.nu / cfg.pr[0] * kx ** 2 * t[J] - 1j * kx * (alpha * t[J] + beta0 * a[J])
                              - gp(pt, qt), 1e-12))
    return worst


# ---------------------------------------------------------------------------
# horizontal products of a y-independent flow against numpy's FFT
# ---------------------------------------------------------------------------
def _spectrum_to_grid(cfg, coeff):
    """Kept slots (K packed, I) -> physical (NZ, NX) through numpy's FFT."""
    nz, nx = cfg.nz, cfg.nx
    S = np.zeros((nz, nx // 2 + 1), complex)
    for k in range(cfg.tnkz + 1):
        s = k if k <= cfg.nkz else k - (cfg.tnkz + 1)
        S[s % nz, :cfg.nkx + 1] = coeff[k, :cfg.nkx + 1]
    return np.fft.irfft2(S, s=(nz, nx)) * (nx * nz)


def _grid_to_kept(cfg, f):
    nz, nx = cfg.nz, cfg.nx
    Fh = np.fft.rfft2(f) / (nx * nz)
    out = np.zeros((cfg.tnkz + 1, cfg.nkx + 1), complex)
    for k in range(cfg.tnkz + 1):
        s = k if k <= cfg.nkz else k - (cfg.tnkz + 1)
        out[k] = Fh[s % nz, :cfg.nkx + 1]
    return out


def pin_horizontal_products(make):
    """A y-independent flow with V = 0 and random kept modes of U1, U3, TH:
    CF1 = -NU k^2 U1 - d/dx(U1 U1) - d/dz(U1 U3), CF3 = -NU k^2 U3
    - d/dx(U1 U3) - d/dz(U3 U3), CFTH = -(NU/PR) k^2 TH - d/dx(U1 TH)
    - d/dz(U3 TH) (for/channel.f:175-189, 410-465, 670-697), with the products
    formed on the NX x NZ grid and truncated to the kept slots; evaluated
    with numpy.fft (an independent library), rows 2..NY."""
    cfg = case.channel_case(24, 17, 18, kick=0.0).replace(nu=0.013, pr=[0.9])
    o = make(cfg)
    rng = np.random.default_rng(5)
    shp = (cfg.tnkz + 1, cfg.nkx + 1)
    fields = {}
    for fid in (U1, U3, CTH0):
        m = crand(rng, shp[0] * shp[1]).reshape(shp) / (1 + np.arange(shp[1]))[None, :]
        m[0, 0] = m[0, 0].real
        for k in range(1, cfg.nkz + 1):
            m[conj_slot(cfg, k), 0] = np.conj(m[k, 0])
        fields[fid] = m
        arr = fv(o, fid)
        arr[:, :shp[0], :shp[1]] = m[None]
    o.rk_chan_1(1)
    u, w, th = (_spectrum_to_grid(cfg, fields[f]) for f in (U1, U3, CTH0))
    kx = np.array([kx_of(cfg, i) for i in range(shp[1])])[None, :]
    kz = np.array([kz_of(cfg, k) for k in range(shp[0])])[:, None]
    k2 = kx ** 2 + kz ** 2
    uu, uw, ww = (_grid_to_kept(cfg, x) for x in (u * u, u * w, w * w))
    ut, wt = _grid_to_kept(cfg, u * th), _grid_to_kept(cfg, w * th)
    want = {F1: -cfg.nu * k2 * fields[U1] - 1j * kx * uu - 1j * kz * uw,
            F3: -cfg.nu * k2 * fields[U3] - 1j * kx * uw - 1j * kz * ww,
            CFTH: -cfg.nu / cfg.pr[0] * k2 * fields[CTH0] - 1j * kx * ut - 1j * kz * wt}
    worst = 0.0
    for fid, w_ in want.items():
        got = fv(o, fid)[2:cfg.ny + 1, :shp[0], :shp[1]]
        worst = max(worst, expect(f"2-D products field {fid}", got, np.broadcast_to(w_, got.shape), 1e-13))
    return worst


# ---------------------------------------------------------------------------
# conservation of the plane means between free-slip walls
# ---------------------------------------------------------------------------
def pin_free_slip_conservation(make):
    """The Kelvin-Helmholtz deck (U, W, TH Neumann, V Dirichlet) with seeded
    noise of amplitude 0.05.  After one step V(2) = V(NY) = 0 (the V rows of
    the solve and the Neumann rows of the projection), and from then on the
    mean modes change only by flux differences that telescope (the face
    fluxes of for/channel.f:510-550, 703-720 and the viscous rows), so
    sum_{J=2..NY-1} DYF(J) <U1>, <U3> and sum over the scalar's rows of
    DYF(J) <TH> stay to round-off for five more steps."""
    cfg = case.kh_case(24, 33, 24, kick=0.05, dt=0.01)
    o = make(cfg)
    o.initialize(*case.initial_physical(cfg))
    o.step()
    dyf = o.vec(DYF)
    ny = cfg.ny
    r = slice(2, ny)

    def sums():
        return [float(np.sum(dyf[r] * fv(o, f)[r, 0, 0].real)) for f in (U1, U3, CTH0)]

    def scales():
        return [float(np.sum(dyf[r] * np.abs(fv(o, f)[r, 0, 0].real))) + 1e-300 for f in (U1, U3, CTH0)]

    s0, sc = sums(), scales()
    for _ in range(5):
        o.step()
    s1 = sums()
    worst = 0.0
    for name, a, b, scl in zip(("U1", "U3", "TH"), s0, s1, sc):
        e = abs(b - a) / scl
        worst = max(worst, e)
        check(e <= 2e-14, f"free-slip mean of {name} drifted by {e:.2e} (relative)")
    check(np.abs(fv(o, U2)[[2, ny], :cfg.tnkz + 1, :cfg.nxp]).max() < 1e-14, "V(2), V(NY) must be zero")
    return worst


# ---------------------------------------------------------------------------
# explicit pressure gradient, projection and CP update
# ---------------------------------------------------------------------------
def pin_pressure_projection(make):
    """Zero velocity, negligible viscosity (NU = 1e-30), free-slip U/W and
    V = 0 walls, and a random CP with CP(1) = CP(2), CP(NY-1) = CP(NY) in
    every mode.  The explicit step puts U* = -H_BAR grad CP (for/channel.f:
    124-148, both wall rows of every solve agree with it), REM_DIV_CHAN's
    discrete Laplacian is the divergence of that gradient, so phi =
    -H_BAR (CP - gauge) exactly: the velocity returns to zero and the CP
    update CP += phi / H_BAR (for/channel.f:1059-1065, rows 2..NY) leaves 0
    in every mode but (0,0), which keeps CP(1) (the phi(1) = 0 gauge);
    CP(1) itself is never updated.  All three RK stages (H_BAR differs from
    BETA_BAR H_BAR at stages 2 and 3)."""
    worst = 0.0
    base = case.kh_case(24, 33, 18, ri=0.0, kick=0.0, dt=0.05).replace(nu=1e-30)
    for rk in (1, 2, 3):
        o = make(base)
        rng = np.random.default_rng(20 + rk)
        ny = base.ny
        cp = fv(o, P)
        shp = (ny + 2, base.tnkz + 1, base.nkx + 1)
        m = crand(rng, np.prod(shp)).reshape(shp)
        m[:, 0, 0] = m[:, 0, 0].real
        for k in range(1, base.nkz + 1):
            m[:, conj_slot(base, k), 0] = np.conj(m[:, k, 0])
        m[1] = m[2]
        m[ny - 1] = m[ny]
        cp[:, :shp[1], :shp[2]] = m
        o.rk_chan_1(rk)
        h = RK_H[rk - 1] * base.delta_t
        kmax = math.hypot(kx_of(base, base.nkx), kz_of(base, base.nkz))
        vscale = h * kmax * np.abs(m).max()
        for f in (U1, U2, U3):
            e = np.abs(fv(o, f)[:, :shp[1], :shp[2]]).max() / vscale
            worst = max(worst, e)
            check(e <= 1e-13, f"RK {rk}: velocity {f} left by the pressure projection: {e:.2e}")
        want = np.zeros(shp, complex)
        want[1] = m[1]
        want[0] = m[0]
        want[ny + 1] = m[ny + 1]
        want[2:ny + 1, 0, 0] = m[1, 0, 0]
        worst = max(worst, expect(f"RK {rk}: CP after the update", fv(o, P)[:, :shp[1], :shp[2]], want, 1e-12))
    return worst


# ---------------------------------------------------------------------------
# mean forcing: the laminar profile (F_TYPE = 1 and 2)
# ---------------------------------------------------------------------------
def _d_yy(o, ny):
    """D_yy on GYF rows 2..NY-1 with rows 1 and NY held at zero."""
    dy, dyf = o.vec(DY), o.vec(DYF)
    n = ny - 2
    M = np.zeros((n, n))
    for r, j in enumerate(range(2, ny)):
        lo, hi = 1.0 / (dy[j] * dyf[j]), 1.0 / (dy[j + 1] * dyf[j])
        M[r, r] = -(lo + hi)
        if r > 0:
            M[r, r - 1] = lo
        if r < n - 1:
            M[r, r + 1] = hi
    return M


def _laminar_fixed_point(make, cfg, time):
    o = make(cfg)
    target = cfg.px0 if cfg.f_type == 1 else cfg.px0 + cfg.amp_omega0 * math.cos(cfg.omega0 * time)
    M = _d_yy(o, cfg.ny)
    u = np.linalg.solve(cfg.nu * M, np.full(cfg.ny - 2, target))
    fv(o, U1)[2:cfg.ny, 0, 0] = u
    o.time = time
    o.step()
    return expect(f"laminar profile F_TYPE={cfg.f_type}", fv(o, U1)[2:cfg.ny, 0, 0].real, u, 1e-13)


def pin_mean_forcing(make):
    """The discrete laminar profile NU D_yy U = -dP/dx is a fixed point:
    F_TYPE = 1 with PX0 (for/channel.f:155-158), and F_TYPE = 2 with
    PX0 + AMP_OMEGA0 cos(OMEGA0 TIME) at the step's TIME (for/channel.f:
    159-164), where PX0 is shifted so that the total equals the F_TYPE = 1
    gradient at that TIME."""
    cfg = case.channel_case(24, 33, 18, kick=0.0)
    e1 = _laminar_fixed_point(make, cfg, 0.0)
    t0, om, amp = 0.83, 2.3, 0.6 * abs(cfg.px0)
    cfg2 = cfg.replace(f_type=2, omega0=om, amp_omega0=amp, px0=cfg.px0 - amp * math.cos(om * t0))
    e2 = _laminar_fixed_point(make, cfg2, t0)
    return max(e1, e2)


# ---------------------------------------------------------------------------
# Crank-Nicolson halves: discrete eigenvectors
# ---------------------------------------------------------------------------
def _cn_factor(mu, h):
    return (1 + 0.5 * h * mu) / (1 - 0.5 * h * mu)


def pin_cn_mean_velocity(make):
    """Eigenvectors of D_yy (Dirichlet walls) in the mean modes of U1 and U3
    decay by prod_k (1 + a_k mu)/(1 - a_k mu), a_k = NU NU_V_SCALE H_BAR(k)/2
    (the explicit half of for/channel.f:606-617 and the implicit rows); no
    other term acts on them."""
    cfg = case.channel_case(16, 25, 16, kick=0.0, dt=0.02).replace(px0=0.0, f_type=3, nu=0.05, nu_v_scale=1.3)
    o = make(cfg)
    M = _d_yy(o, cfg.ny)
    mu, vecs = np.linalg.eig(M)
    order = np.argsort(-mu.real)
    v1, v3 = vecs[:, order[0]].real, vecs[:, order[3]].real
    m1, m3 = mu[order[0]].real, mu[order[3]].real
    fv(o, U1)[2:cfg.ny, 0, 0] = v1
    fv(o, U3)[2:cfg.ny, 0, 0] = v3
    o.step()
    f1 = np.prod([_cn_factor(cfg.nu * cfg.nu_v_scale * m1, h * cfg.delta_t) for h in RK_H])
    f3 = np.prod([_cn_factor(cfg.nu * cfg.nu_v_scale * m3, h * cfg.delta_t) for h in RK_H])
    check(f1 < 0.9999 and f3 < 0.999, "the decay must be resolved")
    e1 = expect("U1 mean CN decay", fv(o, U1)[2:cfg.ny, 0, 0].real, f1 * v1, 1e-12)
    e3 = expect("U3 mean CN decay", fv(o, U3)[2:cfg.ny, 0, 0].real, f3 * v3, 1e-12)
    return max(e1, e3)


def pin_cn_scalar(make):
    """The scalar's mean mode, eigenvector of D_yy with Dirichlet walls,
    decays by the CN factors with a_k = NU NU_V_SCALE H_BAR(k) / (2 PR)
    (for/channel.f:740-746 and the scalar's implicit rows)."""
    cfg = case.channel_case(16, 25, 16, kick=0.0, dt=0.02).replace(
        px0=0.0, f_type=3, th_bc_ymax_c1=[0.0], pr=[0.7], nu=0.05, nu_v_scale=1.3)
    o = make(cfg)
    M = _d_yy(o, cfg.ny)
    mu, vecs = np.linalg.eig(M)
    order = np.argsort(-mu.real)
    v, m = vecs[:, order[2]].real, mu[order[2]].real
    fv(o, CTH0)[2:cfg.ny, 0, 0] = v
    o.step()
    f = np.prod([_cn_factor(cfg.nu * cfg.nu_v_scale / cfg.pr[0] * m, h * cfg.delta_t) for h in RK_H])
    check(f < 0.999, "the decay must be resolved")
    return expect("TH mean CN decay", fv(o, CTH0)[2:cfg.ny, 0, 0].real, f * v, 1e-12)


def pin_cn_stokes_mode(make):
    """A divergence-free (u, v) pair in mode (5, 0) between free-slip walls
    (V = 0 at rows 1, 2, NY, NY+1; U Neumann): v an eigenvector of the V
    operator ((v(J+1)-v(J))/DYF(J) - (v(J)-v(J-1))/DYF(J-1))/DY(J) on rows
    3..NY-1 with eigenvalue lam, u(J) = i (v(J+1)-v(J)) / (kx DYF(J)) on
    rows 2..NY-1 (continuity) and u(1) = u(2), u(NY) = u(NY-1).  u is then
    an eigenvector of the U operator with the same lam, so the CN halves
    (for/channel.f:606-626) and the RK3 horizontal viscosity -NU kx^2 (with
    its ZETA_BAR carry-over) scale both by the same factor per substep, the
    projection has nothing to remove, and after one step both are the
    scalar recurrence's factor times their start (amplitude 1e-8: the
    nonlinear terms are 1e-8 relative and fall at 0 and 10, outside the
    kept band, or in the mean mode)."""
    cfg = case.kh_case(24, 33, 18, ri=0.0, kick=0.0, dt=0.05).replace(nu=0.05, nu_v_scale=1.3)
    o = make(cfg)
    ny = cfg.ny
    dy, dyf = o.vec(DY), o.vec(DYF)
    rows = list(range(3, ny))
    n = len(rows)
    A = np.zeros((n, n))
    for r, j in enumerate(rows):
        lo, hi = 1.0 / (dyf[j - 1] * dy[j]), 1.0 / (dyf[j] * dy[j])
        A[r, r] = -(lo + hi)
        if r > 0:
            A[r, r - 1] = lo
        if r < n - 1:
            A[r, r + 1] = hi
    lam_all, vecs = np.linalg.eig(A)
    order = np.argsort(-lam_all.real)
    lam = lam_all[order[2]].real
    v = np.zeros(ny + 2, complex)
    v[3:ny] = vecs[:, order[2]].real * (1e-8 * (0.6 + 0.8j))
    kx = kx_of(cfg, 5)
    u = np.zeros(ny + 2, complex)
    for j in range(2, ny):
        u[j] = 1j * (v[j + 1] - v[j]) / (kx * dyf[j])
    u[1], u[ny] = u[2], u[ny - 1]
    fv(o, U1)[:, 0, 5] = u
    fv(o, U2)[:, 0, 5] = v
    o.step()
    x, cf_old = 1.0, 0.0
    mu = cfg.nu * cfg.nu_v_scale * lam
    for k in range(3):
        h = RK_H[k] * cfg.delta_t
        cf_new = -cfg.nu * kx ** 2 * x
        x = (x + RK_ZETA[k] * h * cf_old + RK_BETA[k] * h * cf_new + 0.5 * h * mu * x) / (1 - 0.5 * h * mu)
        cf_old = cf_new
    check(abs(x - 1) > 1e-5, "the decay must be resolved")
    e1 = expect("Stokes mode U1", fv(o, U1)[1:ny + 1, 0, 5], x * u[1:ny + 1], 1e-11)
    e2 = expect("Stokes mode U2", fv(o, U2)[1:ny + 2, 0, 5], x * v[1:ny + 2], 1e-11)
    return max(e1, e2)


PINS = {
    "linear_terms": pin_linear_terms,
    "buoyancy": pin_buoyancy,
    "v_products_stretched": pin_v_products_stretched,
    "v_products_uniform": pin_v_products_uniform,
    "horizontal_products": pin_horizontal_products,
    "free_slip_conservation": pin_free_slip_conservation,
    "pressure_projection": pin_pressure_projection,
    "mean_forcing": pin_mean_forcing,
    "cn_mean_velocity": pin_cn_mean_velocity,
    "cn_scalar": pin_cn_scalar,
    "cn_stokes_mode": pin_cn_stokes_mode,
}

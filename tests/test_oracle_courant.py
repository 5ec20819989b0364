"""The oracle's COURANT (variable time step) pinned by brute force over the
physical arrays, by the limits that do not involve the velocity, and by where
RK_CHAN_1 calls it: after the implicit solves of RK_STEP 3, before the
projection, with the step itself still on the old DELTA_T."""
import math

import numpy as np
import pytest

from diablo_amd import case


_MAKE = ("nx", "ny", "nz", "n_steps", "dt", "n_th")


def _cfg(**kw):
    base = dict(nx=24, ny=9, nz=18, n_steps=1)
    base.update(kw)
    made = case.channel_case(**{k: v for k, v in base.items() if k in _MAKE})
    return made.replace(**{k: v for k, v in base.items() if k not in _MAKE})


def _zero(o, oracle_mod):
    for f in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3):
        o.field(f)[...] = 0.0


def _brute(cfg, o, oracle_mod, cfl):
    """min over j = 1..NY, k < NZ, i < NX of cfl * spacing / |u| (no other limit)."""
    dy = o.vec(oracle_mod.DY)
    u = [np.abs(o.field(f)[1:cfg.ny + 1, :cfg.nz, :cfg.nx]) for f in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3)]
    with np.errstate(divide="ignore"):
        qx = (cfl * (cfg.lx / cfg.nx)) / u[0]
        qy = (cfl * dy[1:cfg.ny + 1])[:, None, None] / u[1]
        qz = (cfl * (cfg.lz / cfg.nz)) / u[2]
    return float(min(qx.min(), qy.min(), qz.min()))


def test_viscous_limit_and_caps(oracle_mod):
    """Zero velocity: the diffusive limit 0.5 min(dx, dy(1))^2 / nu, whichever
    of dx and dy(1) is smaller; 999 cap; H_BAR follows DELTA_T."""
    cfg = _cfg(n_th=0)
    o = oracle_mod.Oracle(cfg)
    _zero(o, oracle_mod)
    dy1 = o.vec(oracle_mod.DY)[1]
    dx = cfg.lx / cfg.nx
    assert dy1 < dx
    assert o.courant(0.5) == pytest.approx(0.5 * dy1 * dy1 / cfg.nu, rel=1e-15)
    assert o.courant(7.0) == o.courant(0.5)            # no velocity: cfl plays no part
    o2 = oracle_mod.Oracle(cfg.replace(nu=1.0e-9))
    _zero(o2, oracle_mod)
    assert o2.courant(0.5) == 999.0                     # capped
    # coarse y grid: dx is the smaller spacing
    cfg3 = _cfg(n_th=0, nx=8, ny=5, nz=8, lx=0.5)
    o3 = oracle_mod.Oracle(cfg3)
    _zero(o3, oracle_mod)
    assert o3.vec(oracle_mod.DY)[1] > cfg3.lx / cfg3.nx
    assert o3.courant(0.5) == pytest.approx(0.5 * (cfg3.lx / cfg3.nx) ** 2 / cfg3.nu, rel=1e-15)


def test_scalar_rotation_and_buoyancy_limits(oracle_mod):
    cfg = _cfg(n_th=1, nu=1.0e-5)
    base = oracle_mod.Oracle(cfg)
    _zero(base, oracle_mod)
    d0 = base.courant(0.5)
    # PR < 1 scales the diffusive limit by PR, PR >= 1 leaves it
    for pr, want in ((0.25, 0.25 * d0), (4.0, d0)):
        o = oracle_mod.Oracle(cfg.replace(pr=[pr]))
        _zero(o, oracle_mod)
        assert o.courant(0.5) == pytest.approx(want, rel=1e-15)
    # rotation: a twentieth of the inertial period 2 pi / |I_RO|
    o = oracle_mod.Oracle(cfg.replace(i_ro=-50.0))
    _zero(o, oracle_mod)
    assert o.courant(0.5) == pytest.approx(2 * math.pi / 50.0 / 20.0, rel=1e-15)
    # buoyancy: 0.1 * 2 pi / N with N^2 = |RI * dTH/dy at the wall|; the
    # reference's 0.1 is a single-precision literal, so the factor is
    # float32(0.1) = 0.100000001490116..., not 0.1
    o = oracle_mod.Oracle(cfg.replace(ri=[4.0], th_bc_ymin_c1=[-9.0]))
    _zero(o, oracle_mod)
    got = o.courant(0.5)
    assert got == float(np.float32(0.1)) * 2.0 * math.pi / 6.0
    assert got != 0.1 * 2.0 * math.pi / 6.0


def test_velocity_limit_brute_force(oracle_mod):
    """Random physical velocity, with large values planted where COURANT must
    not look (the x and z pads, planes 0 and NY+1)."""
    cfg = _cfg(n_th=0, nu=1.0e-6)
    o = oracle_mod.Oracle(cfg)
    rng = np.random.default_rng(5)
    for f in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3):
        a = o.field(f)
        a[...] = rng.standard_normal(a.shape)
        a[:, :, cfg.nx:] = 1.0e6
        a[:, cfg.nz:, :] = 1.0e6
        a[0] = 1.0e6
        a[cfg.ny + 1] = 1.0e6
    want = _brute(cfg, o, oracle_mod, 0.7)
    assert want < 0.5 * o.vec(oracle_mod.DY)[1] ** 2 / cfg.nu
    assert o.courant(0.7) == want
    assert o.delta_t == want


@pytest.mark.parametrize("which", [0, 1, 2])
def test_velocity_limit_picks_spacing(oracle_mod, which):
    """One large value of one component at plane j: the limit is cfl times the
    spacing of that direction at that plane (DY(j), not DY(j+1)) over it."""
    cfg = _cfg(n_th=0, nu=1.0e-6)
    o = oracle_mod.Oracle(cfg)
    _zero(o, oracle_mod)
    j, k, i = 4, 3, 5
    f = (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3)[which]
    o.field(f)[j, k, i] = -40.0
    dy = o.vec(oracle_mod.DY)
    spacing = (cfg.lx / cfg.nx, dy[j], cfg.lz / cfg.nz)[which]
    assert o.courant(0.3) == (0.3 * spacing) / 40.0


def test_negative_dt_falls_back(oracle_mod):
    cfg = _cfg(n_th=0)
    o = oracle_mod.Oracle(cfg)
    _zero(o, oracle_mod)
    o.field(oracle_mod.U1)[2, 1, 1] = 1.0
    assert o.courant(-0.5) == 0.0001


def _shear_state(cfg, oracle_mod):
    """Mean flow U(y) plus w(x, y): divergence free, and it stays so through
    the step (the only nonlinear term is U dw/dx), so the projection leaves the
    velocity COURANT sees unchanged up to round-off."""
    o = oracle_mod.Oracle(cfg)
    for fid in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3, oracle_mod.P):
        o.field(fid)[...] = 0.0
    gyf = o.vec(oracle_mod.GYF)
    c1 = case.fourier_view(o.field(oracle_mod.U1))
    c3 = case.fourier_view(o.field(oracle_mod.U3))
    y = gyf[:]
    c1[:, 0, 0] = 3.0 * y * (2.0 - y)
    c3[:, 0, 1] = (0.4 + 0.3j) * np.sin(np.pi * y / 2.0)
    c3[:, 0, 2] = (0.1 - 0.2j) * np.cos(np.pi * y / 2.0)
    return o


def test_in_step_call_sees_velocity_after_the_solves(oracle_mod):
    """RK_STEP 3 with COURANT: the step's own fields are those of a fixed-dt
    step (TEMP1..TEMP5 were set before), the new DELTA_T is COURANT of the
    velocity after the implicit solves -- here equal to the final velocity --
    and not of the velocity the step started from."""
    cfg = _cfg(n_th=0, nu=1.0e-2, dt=2.0e-2)
    a, b = _shear_state(cfg, oracle_mod), _shear_state(cfg, oracle_mod)
    start = _shear_state(cfg, oracle_mod)
    a.rk_chan_1(1); a.rk_chan_1(2)
    b.rk_chan_1(1); b.rk_chan_1(2)
    for fid in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3):
        start.field(fid)[...] = a.field(fid)
    a.L.dor_courant_request(a.h, 0.4)
    a.rk_chan_1(3)
    b.rk_chan_1(3)
    for fid in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3, oracle_mod.P):
        np.testing.assert_array_equal(a.field(fid), b.field(fid))
    assert b.delta_t == cfg.delta_t
    dt_in_step = a.delta_t
    assert dt_in_step != cfg.delta_t
    # COURANT of the final velocity (physical copies)
    for fid in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3):
        b.fft_xz_to_physical(fid, 0, cfg.ny + 1)
    want = _brute(cfg, b, oracle_mod, 0.4)
    assert dt_in_step == pytest.approx(want, rel=1e-12)
    # ... and not of the velocity at the start of RK_STEP 3
    for fid in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3):
        start.fft_xz_to_physical(fid, 0, cfg.ny + 1)
    assert abs(_brute(cfg, start, oracle_mod, 0.4) / want - 1.0) > 1e-6
    # only RK_STEP 3 calls it
    c = _shear_state(cfg, oracle_mod)
    c.L.dor_courant_request(c.h, 0.4)
    c.rk_chan_1(1)
    assert c.delta_t == cfg.delta_t


def test_step_advances_time_by_new_dt(oracle_mod):
    """TIME += DELTA_T after the step uses the DELTA_T that COURANT set
    (for/diablo.f:109 runs after RK_CHAN_1)."""
    cfg = _cfg(n_th=0, nu=1.0e-2, dt=2.0e-2)
    o = _shear_state(cfg, oracle_mod)
    o.step(courant_cfl=0.4)
    assert o.time == o.delta_t != cfg.delta_t

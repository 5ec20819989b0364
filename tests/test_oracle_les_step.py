"""The oracle's RK_CHAN_1 with LES on, pinned by closed forms of the discrete
equations solved independently in numpy:

* with the eddy viscosity NU_T(y) held fixed (les on, the model not called),
  an eigenvector mu of the (0,0)-mode operator
  nu D_yy + d/dy(NU_T d/dy) (flux form, NU_T on the GY faces) decays per step
  by prod_k (1 + h_k mu / 2) / (1 - h_k mu / 2): the explicit and implicit
  halves of the eddy-viscosity Crank-Nicolson terms agree with each other and
  with the face placement of NU_T; the same for a scalar with KAPPA_T;
* NU_T = 0 gives the DNS step bit for bit;
* with the model called, the laminar profile (no fluctuations, so |S| with
  the mean shear removed vanishes up to transform round-off) stays a fixed
  point."""
import numpy as np
import pytest

from diablo_amd import case


def _zero_state(o, oracle_mod, cfg):
    for f in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3, oracle_mod.P):
        o.field(f)[...] = 0.0
    for n in range(cfg.n_th):
        o.field(oracle_mod.CTH0 + n)[...] = 0.0


def _operator(o, oracle_mod, cfg, diff, nt):
    """diff * D_yy + d/dy(nt d/dy) on GYF rows 2..NY-1, rows 1 and NY held at
    zero: the flux through the GY face j is nt(j) (u(j) - u(j-1)) / DY(j)."""
    ny = cfg.ny
    dy, dyf = o.vec(oracle_mod.DY), o.vec(oracle_mod.DYF)
    n = ny - 2
    M = np.zeros((n, n))
    for r, j in enumerate(range(2, ny)):
        lo = (diff + nt[j]) / (dy[j] * dyf[j])
        hi = (diff + nt[j + 1]) / (dy[j + 1] * dyf[j])
        M[r, r] = -(lo + hi)
        if r > 0:
            M[r, r - 1] = lo
        if r < n - 1:
            M[r, r + 1] = hi
    return M


def _factor(cfg, mu):
    fac = 1.0
    for h in np.array([8.0, 2.0, 5.0]) / 15.0 * cfg.delta_t:
        fac *= (1 + h * mu / 2) / (1 - h * mu / 2)
    return fac


@pytest.mark.parametrize("pick", [0, 2])
def test_eddy_viscosity_cn_decay(oracle_mod, pick):
    cfg = case.channel_case(16, 25, 16, kick=0.0, dt=0.02, n_steps=1).replace(
        px0=0.0, th_bc_ymax_c1=[0.0], pr=[0.7], nu=0.02)
    ny = cfg.ny
    gy = case.grid_gy(cfg)
    nt = np.zeros(ny + 2)
    nt[1:ny + 2] = 0.03 * (1.2 + np.sin(3.0 * gy))        # NU_T(GY(j)), j = 1..NY+1
    kt = np.zeros(ny + 2)
    kt[1:ny + 2] = 0.05 * (1.1 + np.cos(2.0 * gy))
    o = oracle_mod.Oracle(cfg)
    _zero_state(o, oracle_mod, cfg)
    o.set_les(True, model=False)
    o.field(oracle_mod.NUT)[...] = nt[:, None, None]
    o.field(oracle_mod.KAPPA0)[...] = kt[:, None, None]
    Mu = _operator(o, oracle_mod, cfg, cfg.nu, nt)
    Mt = _operator(o, oracle_mod, cfg, cfg.nu / cfg.pr[0], kt)
    mu_u, vu = np.linalg.eig(Mu)
    mu_t, vt = np.linalg.eig(Mt)
    iu, it = np.argsort(-mu_u.real)[pick], np.argsort(-mu_t.real)[pick]
    u0, t0 = vu[:, iu].real, vt[:, it].real
    case.fourier_view(o.field(oracle_mod.U1))[2:ny, 0, 0] = u0
    case.fourier_view(o.field(oracle_mod.CTH0))[2:ny, 0, 0] = t0
    o.step()
    fu, ft = _factor(cfg, mu_u[iu].real), _factor(cfg, mu_t[it].real)
    gu = case.fourier_view(o.field(oracle_mod.U1))[2:ny, 0, 0]
    gt = case.fourier_view(o.field(oracle_mod.CTH0))[2:ny, 0, 0]
    assert np.abs(gu - fu * u0).max() < 1e-12 * np.abs(u0).max(), (fu,)
    assert np.abs(gt - ft * t0).max() < 1e-12 * np.abs(t0).max(), (ft,)
    # the eddy terms matter at this bar: the molecular operator alone decays differently
    fm = _factor(cfg, np.linalg.eigvals(_operator(o, oracle_mod, cfg, cfg.nu, 0 * nt)).real.max())
    assert abs(fu - fm) > 1e-3


def test_zero_eddy_viscosity_is_the_dns_step(oracle_mod):
    cfg = case.channel_case(24, 25, 18, kick=0.3, streak=0.1)
    a = oracle_mod.Oracle(cfg)
    a.initialize(*case.initial_physical(cfg))
    b = oracle_mod.Oracle(cfg)
    b.initialize(*case.initial_physical(cfg))
    b.set_les(True, model=False)
    for _ in range(2):
        a.step()
        b.step()
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


def test_laminar_profile_fixed_point_with_model(oracle_mod):
    cfg = case.channel_case(24, 33, 18, kick=0.0, n_steps=1)
    o = oracle_mod.Oracle(cfg)
    _zero_state(o, oracle_mod, cfg)
    ny = cfg.ny
    M = _operator(o, oracle_mod, cfg, 1.0, np.zeros(ny + 2))
    u = np.linalg.solve(cfg.nu * M, np.full(ny - 2, cfg.px0))
    case.fourier_view(o.field(oracle_mod.U1))[2:ny, 0, 0] = u
    o.set_les(True, model=True)
    for _ in range(3):
        o.step()
    assert np.abs(o.field(oracle_mod.NUT)).max() < 1e-14 * cfg.nu     # transform round-off only
    got = case.fourier_view(o.field(oracle_mod.U1))
    assert np.abs(got[2:ny, 0, 0] - u).max() < 1e-13 * np.abs(u).max()
    k = case.kept(cfg, o.field(oracle_mod.U1)).copy()
    k[:, 0, 0] = 0.0
    assert np.abs(k).max() < 1e-15


def test_tke_les_summation_by_parts(oracle_mod):
    """tkebudget_chan_les on mean profiles with a y-dependent NU_T (zero on
    GHOST_LES's rows): summed over the column with the staggered weights, the
    SGS work of each velocity component equals minus the eddy dissipation
    of its discrete gradient (summation by parts), which fixes the face
    placement of NU_T in both stencils."""
    cfg = case.channel_case(16, 25, 16, kick=0.0)
    ny, n = cfg.ny, cfg.nx * cfg.nz
    o = oracle_mod.Oracle(cfg)
    _zero_state(o, oracle_mod, cfg)
    o.set_les(True, model=False)
    gy = np.concatenate([[0.0], case.grid_gy(cfg)])                    # GY(0..NY+1), GY(0) unused
    nt = np.zeros(ny + 2)
    nt[3:ny] = 0.02 * (1.5 + np.sin(2.0 * gy[3:ny]))
    o.field(oracle_mod.NUT)[...] = nt[:, None, None]
    dy, dyf = o.vec(oracle_mod.DY), o.vec(oracle_mod.DYF)
    rng = np.random.default_rng(3)
    u = rng.standard_normal(ny + 2)
    case.fourier_view(o.field(oracle_mod.U1))[:, 0, 0] = u
    eps = o.tke_les()
    # plane 1: the reference's U2 term divides 0 by DYF(0), which it never
    # sets, so epsilon_sgs(1) is NaN there too; its exact value is 0 (NU_T
    # vanishes on the faces of plane 1), so the sums start at plane 2
    assert np.isnan(eps[0])
    lhs = float(np.sum(dyf[2:ny + 1] * eps[1:]))
    rhs = -n * float(np.sum(nt[2:ny + 1] * (u[2:ny + 1] - u[1:ny]) ** 2 / dy[2:ny + 1]))
    assert rhs < 0 and abs(lhs - rhs) < 1e-12 * abs(rhs), (lhs, rhs)
    # the wall-normal component on GY rows, NU_T averaged onto the GYF faces
    case.fourier_view(o.field(oracle_mod.U1))[...] = 0.0
    case.fourier_view(o.field(oracle_mod.U2))[:, 0, 0] = u
    eps = o.tke_les()
    lhs = float(np.sum(dy[2:ny + 1] * eps[1:]))
    nf = 0.5 * (nt[1:ny] + nt[2:ny + 1])                                 # faces GYF(1..NY-1)
    rhs = -n * float(np.sum(nf * (u[2:ny + 1] - u[1:ny]) ** 2 / dyf[1:ny]))
    assert rhs < 0 and abs(lhs - rhs) < 1e-12 * abs(rhs), (lhs, rhs)

"""Shared helpers for the GPU parity tests (compare the device path with the
CPU oracle on the same seeded inputs)."""
import numpy as np

from diablo_amd import case

FIELDS = ("CU1", "CU2", "CU3", "CP", "CF1", "CF2", "CF3", "CTH0", "CFTH0")


def rel_err(cfg, a, b):
    """max |a-b| / max |b| over the kept Fourier slots of all planes."""
    ka, kb = case.kept(cfg, a), case.kept(cfg, b)
    scale = max(np.abs(kb).max(), 1e-300)
    return float(np.abs(ka - kb).max() / scale)


def compare_states(cfg, sg, so, tol, fields=FIELDS):
    errs = {}
    for k in fields:
        if k not in so:
            continue
        errs[k] = rel_err(cfg, sg[k], so[k])
    bad = {k: v for k, v in errs.items() if not v <= tol}
    return errs, bad


def kitchen_sink(nx=24, ny=17, nz=18):
    """A channel deck that switches on every term of RK_CHAN_1: Coriolis,
    buoyancy in all three directions, background scalar gradients, an
    oscillating pressure gradient, anisotropic vertical viscosity, a moving
    wall and Neumann velocity/scalar conditions."""
    cfg = case.channel_case(nx=nx, ny=ny, nz=nz, re=200.0, dt=2e-3, kick=0.05, streak=0.05, seed=7)
    return cfg.replace(i_ro=0.3, grav_x=0.1, grav_y=1.0, grav_z=0.2, ri=[0.5], drhodx=[0.05],
                       drhodz=[-0.02], f_type=2, px0=-0.01, amp_omega0=0.005, omega0=3.0,
                       nu_v_scale=1.3, u_bc_ymax_c1=0.3, w_bc_ymin=1, w_bc_ymin_c1=0.1,
                       th_bc_ymin=[1], th_bc_ymin_c1=[0.2], th_bc_ymax=[0], th_bc_ymax_c1=[1.0])

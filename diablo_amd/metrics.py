"""Error and energy diagnostics used to state the mixed path's accuracy
against the FP64 path (relative L2 velocity error, kinetic-energy drift,
max |div u|).  All sums are FP64 on host arrays; the physical fields come from
the library's own FFT_XZ_TO_PHYSICAL."""
from __future__ import annotations

import numpy as np

from . import case


def physical_velocity(g, state: dict):
    """u, v, w in physical space (planes 0..NY+1, NZ, NX) from a state dict."""
    cfg = g.cfg
    out = []
    for k in ("CU1", "CU2", "CU3"):
        a = np.array(state[k], copy=True)
        g.fft_xz_to_physical(a, 0, cfg.ny + 1)
        out.append(a[:, :cfg.nz, :cfg.nx].copy())
    return out


def rel_l2(vel, ref, ny: int) -> float:
    """||u - u_ref||_2 / ||u_ref||_2 over u, v, w on rows J = 1..NY."""
    num = sum(float(((a[1:ny + 1] - b[1:ny + 1]) ** 2).sum()) for a, b in zip(vel, ref))
    den = sum(float((b[1:ny + 1] ** 2).sum()) for b in ref)
    return (num / den) ** 0.5


def kinetic_energy(cfg, vel) -> float:
    """sum(u^2 DYF + w^2 DYF + v^2 DY) / (NX NZ LY) over J = 1..NY (u, w live at
    GYF, v at GY; compare INTEGRATE_Y_VAR, for/mpi.f:1676-1698)."""
    _, _, dy, dyf = case.grid_metrics(case.grid_gy(cfg), cfg.ny)
    u, v, w = vel
    j = np.arange(1, cfg.ny + 1)
    e = ((u[j] ** 2).sum(axis=(1, 2)) * dyf[j] + (w[j] ** 2).sum(axis=(1, 2)) * dyf[j]
         + (v[j] ** 2).sum(axis=(1, 2)) * dy[j]).sum()
    return float(e / (cfg.nx * cfg.nz * cfg.ly))


def max_divergence(g, state: dict) -> float:
    """max |div u| in physical space with REM_DIV_CHAN's discrete divergence
    (for/channel.f:1139-1141) on rows J = 2..NY-1."""
    cfg = g.cfg
    kx, kz, _, _ = case.wavenumbers(cfg)
    _, _, _, dyf = case.grid_metrics(case.grid_gy(cfg), cfg.ny)
    k1, k2, k3 = (case.kept(cfg, state[k]) for k in ("CU1", "CU2", "CU3"))
    div = np.zeros(cfg.field_shape())
    dv = case.fourier_view(div)
    j = np.arange(2, cfg.ny)
    dv[j, :cfg.tnkz + 1, :cfg.nkx + 1] = (1j * kx[:cfg.nkx + 1][None, None, :] * k1[j]
                                          + (k2[j + 1] - k2[j]) / dyf[j][:, None, None]
                                          + 1j * kz[None, :, None] * k3[j])
    g.fft_xz_to_physical(div, 0, cfg.ny + 1)
    return float(np.abs(div[2:cfg.ny, :cfg.nz, :cfg.nx]).max())


def compare(g_ref, s_ref: dict, g, s: dict) -> dict:
    """Errors of state s against the reference state s_ref (same case)."""
    cfg = g.cfg
    vr = physical_velocity(g_ref, s_ref)
    v = physical_velocity(g, s)
    e_ref = kinetic_energy(cfg, vr)
    e = kinetic_energy(cfg, v)
    return {"rel_l2": rel_l2(v, vr, cfg.ny), "ke_drift": abs(e - e_ref) / e_ref,
            "max_div": max_divergence(g, s), "max_div_ref": max_divergence(g_ref, s_ref)}

"""Grid generator, grid metrics, wavenumbers, index ranges and the RKW3 table
(create_grid_h5.m:28-75, for/hdf5.f:676-703, for/channel.f:1572-1652,
for/fft.f:72-105, for/mpi.f:1080-1100, for/diablo_io.f:116-124)."""
import math
import os

import numpy as np
import pytest

from diablo_amd import case

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "kh_grid_y.txt")


def test_kh_grid_fixture():
    """create_grid_h5.m restated reproduces the reference's KH_test/grid.h5."""
    ref = np.loadtxt(GOLDEN)
    gy = case.create_grid_gy(2, 65, 20.0, 1.5)
    assert gy.shape == ref.shape == (66,)
    assert np.abs(gy - ref).max() < 4e-15


def test_kh_walls_at_pm10():
    gy = np.loadtxt(GOLDEN)
    _, gyf, _, _ = case.grid_metrics(gy, 65)
    assert abs(gyf[1] + 10.0) < 1e-13 and abs(gyf[65] - 10.0) < 1e-13


@pytest.mark.parametrize("gt,n,L,cs", [(1, 33, 2.0, 1.75), (3, 65, 2.0, 1.75), (2, 97, 20.0, 1.5)])
def test_grid_walls_and_monotone(gt, n, L, cs):
    gy = case.create_grid_gy(gt, n, L, cs)
    _, gyf, dy, dyf = case.grid_metrics(gy, n)
    lo = -L / 2 if gt in (1, 2) else 0.0
    assert abs(gyf[1] - lo) < 1e-13 and abs(gyf[n] - (lo + L)) < 1e-13
    assert np.all(np.diff(gy) > 0)
    assert np.all(dy[1:] > 0) and np.all(dyf[1:] > 0)


def test_grid_metrics_definitions(oracle_mod):
    cfg = case.channel_case(16, 9, 16)
    gy = case.grid_gy(cfg)
    GY, GYF, DY, DYF = case.grid_metrics(gy, cfg.ny)
    o = oracle_mod.Oracle(cfg, gy)
    for vid, ref in ((oracle_mod.GY, GY), (oracle_mod.GYF, GYF), (oracle_mod.DY, DY), (oracle_mod.DYF, DYF)):
        assert np.array_equal(o.vec(vid), ref)
    # closed forms: GY(0) = DY(0) = DYF(0) = 0, ghost GYF by linear extrapolation
    assert GY[0] == 0 and DY[0] == 0 and DYF[0] == 0
    assert GYF[0] == 2 * GYF[1] - GYF[2]
    assert DYF[cfg.ny + 1] == DYF[cfg.ny]
    assert np.allclose(np.cumsum(DYF[1:cfg.ny + 1]), GY[2:cfg.ny + 2] - GY[1])


@pytest.mark.parametrize("nx,nz,lx,lz", [(64, 64, 2 * math.pi, math.pi), (384, 384, 2 * math.pi, math.pi),
                                         (96, 48, 30.0, 4.0), (256, 256, 30.0, 4.0)])
def test_wavenumbers_bit_exact(oracle_mod, nx, nz, lx, lz):
    cfg = case.channel_case(nx, 5, nz).replace(lx=lx, lz=lz)
    kx, kz, kx2, kz2 = case.wavenumbers(cfg)
    pi = 4.0 * math.atan(1.0)
    assert pi == math.pi
    assert kx.size == nx // 2 + 1 and kx[nx // 2] == 0.0
    for i in range(nx // 2):
        assert kx[i] == (i * (2.0 * pi)) / lx
    nkz = nz // 3
    assert kz.size == 2 * nkz + 1
    for k in range(nkz + 1):
        assert kz[k] == (k * (2.0 * pi)) / lz
    for k in range(1, nkz + 1):
        assert kz[2 * nkz + 1 - k] == -kz[k]
    o = oracle_mod.Oracle(cfg)
    assert np.array_equal(o.vec(oracle_mod.KX), kx) and np.array_equal(o.vec(oracle_mod.KZ), kz)
    assert np.array_equal(o.vec(oracle_mod.KX2), kx2) and np.array_equal(o.vec(oracle_mod.KZ2), kz2)


@pytest.mark.parametrize("n,nk", [(64, 21), (96, 32), (128, 42), (192, 64), (256, 85), (384, 128), (16, 5), (24, 8)])
def test_dealias_truncation_counts(n, nk):
    cfg = case.channel_case(n, 5, n)
    assert cfg.nkx == nk and cfg.nkz == nk and cfg.tnkz == 2 * nk
    kz = case.kz_of_slot(cfg)
    assert sorted(kz.tolist()) == list(range(-nk, nk + 1))
    assert kz[cfg.tnkz] == -1 and kz[cfg.nkz + 1] == -nk


def test_index_ranges(oracle_mod):
    for thb, js in ((0, 2), (1, 1)):
        cfg = case.channel_case(16, 9, 16).replace(th_bc_ymin=[thb])
        o = oracle_mod.Oracle(cfg)
        assert o.info(0) == 2 and o.info(1) == cfg.ny
        assert o.info(10) == js and o.info(20) == cfg.ny
        jst, jen, jsth, jenth = case.index_ranges(cfg)
        assert (jst, jen, jsth[0], jenth[0]) == (2, cfg.ny, js, cfg.ny)


def test_rkw3_is_smr_third_order():
    h, b, z = case.rkw3_table(1.0)
    gamma = b * h
    zeta = z * h
    assert np.allclose(gamma, [8 / 15, 5 / 12, 3 / 4]) and np.allclose(zeta, [0, -17 / 60, -5 / 12])
    assert np.isclose(h.sum(), 1.0)
    # third order for a linear ODE: amplification = 1 + z + z^2/2 + z^3/6
    for lam in (-0.3, 0.2, -1.1):
        u_prev, u = 0.0, 1.0
        for k in range(3):
            u, u_prev = u + gamma[k] * lam * u + zeta[k] * lam * u_prev, u
        assert np.isclose(u, 1 + lam + lam ** 2 / 2 + lam ** 3 / 6, rtol=1e-14)

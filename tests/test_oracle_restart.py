"""The oracle's restart-file layout (WriteHDF5 / ReadHDF5 field conversion)
pinned by a closed-form field, the index order of the file array, and the
round trips the G2GF / GF2G pair allows."""
import numpy as np
import pytest

from diablo_amd import case


def _cfg(nx=24, ny=9, nz=18):
    return case.channel_case(nx=nx, ny=ny, nz=nz, n_steps=1)


def _closed_form_state(cfg, o, oracle_mod):
    """u = a(y) + 2 Re(b(y) e^{i(kx x + kz z)}) for one kept (kx, kz) per field."""
    for f in (oracle_mod.U1, oracle_mod.U2, oracle_mod.U3, oracle_mod.P, oracle_mod.CTH0):
        o.field(f)[...] = 0.0
    j = np.arange(cfg.ny + 2)
    modes = {oracle_mod.U1: (1, 2), oracle_mod.U2: (2, 1), oracle_mod.U3: (3, cfg.tnkz),
             oracle_mod.P: (0, 3), oracle_mod.CTH0: (cfg.nkx, cfg.nkz)}
    amp = {}
    for n, (f, (i, k)) in enumerate(modes.items()):
        a = 0.3 + 0.1 * n + 0.05 * j
        b = (0.2 + 0.1j) * np.cos(0.4 * j + n)
        c = case.fourier_view(o.field(f))
        c[:, 0, 0] = a
        c[:, k, i] += b
        if i == 0:                       # kx = 0: the -kz partner holds the conjugate
            kk = cfg.tnkz + 1 - k
            c[:, kk, 0] += np.conj(b)
        amp[f] = (a, b, i, k)
    return amp


def _physical(cfg, o, oracle_mod, amp):
    kx = o.vec(oracle_mod.KX)
    kz = o.vec(oracle_mod.KZ)
    x = np.arange(cfg.nx) * cfg.lx / cfg.nx
    z = np.arange(cfg.nz) * cfg.lz / cfg.nz
    out = {}
    for f, (a, b, i, k) in amp.items():
        ph = np.exp(1j * (kx[i] * x[None, :] + kz[k] * z[:, None]))       # (z, x)
        fac = 2.0 if i > 0 else 2.0                                        # b and its conjugate partner
        out[f] = a[:, None, None] + fac * (b[:, None, None] * ph[None]).real  # (y, z, x)
    return out


def test_file_array_is_the_physical_field(oracle_mod):
    """U, W, P, TH: the file array at (z, y-1, x) (C order) is the physical
    field at (x_i, y_j, z_k) for j = 1..NY; V is its G2GF interpolation."""
    cfg = _cfg()
    o = oracle_mod.Oracle(cfg)
    amp = _closed_form_state(cfg, o, oracle_mod)
    phys = _physical(cfg, o, oracle_mod, amp)
    ny = cfg.ny
    for f in (oracle_mod.U1, oracle_mod.U3, oracle_mod.P, oracle_mod.CTH0):
        got = o.flow_to_file(f)
        want = np.transpose(phys[f][1:ny + 1], (1, 0, 2))                 # (z, y, x)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    v = phys[oracle_mod.U2].copy()
    v[1] = v[2]
    v[ny + 1] = v[ny]
    vf = 0.5 * (v[1:ny + 1] + v[2:ny + 2])
    np.testing.assert_allclose(o.flow_to_file(oracle_mod.U2), np.transpose(vf, (1, 0, 2)), rtol=0, atol=1e-13)
    # writing leaves the state alone
    before = o.field(oracle_mod.U2).copy()
    o.flow_to_file(oracle_mod.U2)
    np.testing.assert_array_equal(o.field(oracle_mod.U2), before)


def test_read_then_write_round_trip(oracle_mod):
    """A band-limited file array comes back unchanged through ReadHDF5's and
    WriteHDF5's conversions.  For V, G2GF takes plane 1 from GF2G's plane 2:
    an array written by WriteHDF5 (plane 1 = V(2)) comes back whole, any
    other loses plane 1 only."""
    cfg = _cfg(32, 13, 24)
    src = oracle_mod.Oracle(cfg)
    amp = _closed_form_state(cfg, src, oracle_mod)
    o = oracle_mod.Oracle(cfg)
    for f in amp:
        arr = src.flow_to_file(f)
        o.flow_from_file(f, arr)
        back = o.flow_to_file(f)
        np.testing.assert_allclose(back, arr, rtol=0, atol=1e-13 * np.abs(arr).max())
    arr = src.flow_to_file(oracle_mod.U1)           # not of WriteHDF5's V form
    o.flow_from_file(oracle_mod.U2, arr)
    back = o.flow_to_file(oracle_mod.U2)
    scale = np.abs(arr).max()
    np.testing.assert_allclose(back[:, 1:, :], arr[:, 1:, :], rtol=0, atol=1e-13 * scale)
    assert np.abs(back[:, 0, :] - arr[:, 0, :]).max() > 1e-6 * scale
    # plane 1 is GF2G's V(2): the alternating sum 2 A(2) - 2 A(3) + ... of the
    # file planes, the last one, A(NY), counted once
    ny = cfg.ny
    sign = np.array([(-1.0) ** (y - 2) for y in range(2, ny + 1)])
    wts = 2.0 * sign
    wts[-1] = sign[-1]
    alt = np.tensordot(wts, np.transpose(arr[:, 1:ny, :], (1, 0, 2)), axes=1)
    np.testing.assert_allclose(back[:, 0, :], alt, rtol=0, atol=1e-12 * scale)


def test_gf2g_inverts_g2gf_and_imposes_walls(oracle_mod):
    """Reading V: planes 2..NY of the Fourier state equal those of the state
    the file came from (GF2G undoes G2GF), and the ghost planes are the
    zero-mass-flux values V(1) = -V(2), V(NY+1) = -V(NY)."""
    cfg = _cfg()
    src = oracle_mod.Oracle(cfg)
    _closed_form_state(cfg, src, oracle_mod)
    o = oracle_mod.Oracle(cfg)
    o.flow_from_file(oracle_mod.U2, src.flow_to_file(oracle_mod.U2))
    a = case.kept(cfg, src.field(oracle_mod.U2))
    b = case.kept(cfg, o.field(oracle_mod.U2))
    ny = cfg.ny
    scale = np.abs(a).max()
    np.testing.assert_allclose(b[2:ny + 1], a[2:ny + 1], rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(b[1], -b[2], rtol=0, atol=1e-15 * scale)
    np.testing.assert_allclose(b[ny + 1], -b[ny], rtol=0, atol=1e-15 * scale)
    assert np.abs(b[0]).max() == 0.0


@pytest.mark.parametrize("fid", ["U1", "CTH0"])
def test_file_index_order(oracle_mod, fid):
    """One nonzero value in the file array lands at (x, y, z) = (i, j, k) of the
    physical field: x fastest, then y (from plane 1), then z."""
    cfg = _cfg(12, 7, 12)
    o = oracle_mod.Oracle(cfg)
    f = getattr(oracle_mod, fid)
    arr = np.zeros((cfg.nz, cfg.ny, cfg.nx))
    arr[5, 2, 3] = 1.0                   # z = 5, y = 3, x = 3
    o.flow_from_file(f, arr)
    w = o.field(f).copy()
    o.fft_array_to_physical(w, 0, cfg.ny + 1)
    phys = w[:, :cfg.nz, :cfg.nx]
    # after dealiasing the spike spreads, but only over plane y = 3, and the
    # plane keeps its sum (the mean mode is kept)
    other = np.delete(phys, 3, axis=0)
    assert np.abs(other).max() < 1e-15
    assert phys[3].sum() == pytest.approx(1.0, rel=1e-12)
    # the kept-mode projection of a spike at (x0, z0) peaks at (x0, z0)
    assert np.unravel_index(np.argmax(phys[3]), phys[3].shape) == (5, 3)

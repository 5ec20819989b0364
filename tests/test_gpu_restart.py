"""Restart-file layout on the device (WriteHDF5 / ReadHDF5 field conversion)
against the oracle, and the file round trip through diablo_amd.restart."""
import numpy as np
import pytest

from diablo_amd import case
from gpu_util import compare_states, kitchen_sink

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu_mod():
    from diablo_amd import gpu
    return gpu


def _pair(oracle_mod, gpu_mod, cfg):
    o = oracle_mod.Oracle(cfg)
    o.initialize(*case.initial_physical(cfg))
    o.step()
    g = gpu_mod.GpuChannel(cfg)
    g.load_state(o.state())
    return o, g


def _ids(oracle_mod, gpu_mod, nth):
    ids = [(oracle_mod.U1, gpu_mod.CU1), (oracle_mod.U2, gpu_mod.CU2), (oracle_mod.U3, gpu_mod.CU3),
           (oracle_mod.P, gpu_mod.CP)]
    ids += [(oracle_mod.CTH0 + n, gpu_mod.CTH0 + n) for n in range(nth)]
    return ids


@pytest.mark.parametrize("maker", ["channel", "sink"])
def test_get_flow_physical(oracle_mod, gpu_mod, maker):
    cfg = {"channel": lambda: case.channel_case(48, 33, 36, kick=0.3, streak=0.1),
           "sink": lambda: kitchen_sink(48, 25, 36)}[maker]()
    say("get_flow_physical", maker)
    o, g = _pair(oracle_mod, gpu_mod, cfg)
    before = g.state()
    for fo, fg in _ids(oracle_mod, gpu_mod, cfg.n_th):
        want = o.flow_to_file(fo)
        got = g.get_flow_physical(fg)
        err = np.abs(got - want).max() / np.abs(want).max()
        say(fg, err)
        assert err < 1e-13, (fg, err)
    after = g.state()
    for k in before:
        assert np.array_equal(before[k], after[k]), k


@pytest.mark.parametrize("maker", ["channel", "sink"])
def test_set_flow_physical(oracle_mod, gpu_mod, maker):
    """The oracle's file arrays read back into a fresh context and into the
    oracle give the same Fourier state."""
    cfg = {"channel": lambda: case.channel_case(48, 33, 36, kick=0.3, streak=0.1),
           "sink": lambda: kitchen_sink(48, 25, 36)}[maker]()
    say("set_flow_physical", maker)
    o, _ = _pair(oracle_mod, gpu_mod, cfg)
    files = {fo: o.flow_to_file(fo) for fo, _ in _ids(oracle_mod, gpu_mod, cfg.n_th)}
    o2 = oracle_mod.Oracle(cfg)
    g2 = gpu_mod.GpuChannel(cfg)
    for fo, fg in _ids(oracle_mod, gpu_mod, cfg.n_th):
        o2.flow_from_file(fo, files[fo])
        g2.set_flow_physical(fg, files[fo])
    errs, bad = compare_states(cfg, g2.state(), o2.state(), 1e-13, fields=("CU1", "CU2", "CU3", "CP", "CTH0"))
    say(errs)
    assert not bad, bad


def test_restart_round_trip_large(gpu_mod, tmp_path):
    """At 192x129x192: save_flow then load_flow into a new context gives the
    state back (V's ghost planes as GF2G sets them), and both contexts step
    on alike."""
    from diablo_amd import restart
    cfg = case.channel_case(192, 129, 192, kick=0.3, streak=0.1)
    say("restart round trip", cfg.nx, cfg.ny, cfg.nz)
    g = gpu_mod.GpuChannel(cfg)
    g.initialize(*case.initial_physical(cfg))
    g.step(2)
    g.synchronize()
    path = str(tmp_path / "flow.npz")
    restart.save_flow(path, g)
    h = gpu_mod.GpuChannel(cfg)
    restart.load_flow(path, h)
    assert h.time == g.time
    # READ_FLOW re-applies the velocity BCs: the ghost planes of the running
    # state hold the pre-projection values, so do the same to it to compare
    g.apply_bc_vel()
    a, b = g.state(), h.state()
    ny = cfg.ny
    for k in ("CU1", "CU3", "CP", "CTH0"):
        ka, kb = case.kept(cfg, a[k]), case.kept(cfg, b[k])
        err = np.abs(ka[1:ny + 1] - kb[1:ny + 1]).max() / np.abs(ka).max()
        say(k, err)
        assert err < 1e-13, (k, err)
    ka, kb = case.kept(cfg, a["CU2"]), case.kept(cfg, b["CU2"])
    assert np.abs(ka[2:ny + 1] - kb[2:ny + 1]).max() < 1e-13 * np.abs(ka).max()
    np.testing.assert_array_equal(kb[1], -kb[2])
    with np.load(path, allow_pickle=False) as f:
        assert tuple(f["Resolution"]) == (cfg.nx, cfg.ny, cfg.nz)
        assert f["U"].shape == (cfg.nz, cfg.ny, cfg.nx)
    # a restart drops CF, which RK_STEP 1 does not read: the velocity goes on
    # alike (the scalar's ghost planes are zero after a restart, in the
    # reference too; it is passive here)
    g.step(1)
    h.step(1)
    errs, bad = compare_states(cfg, h.state(), g.state(), 1e-11, fields=("CU1", "CU2", "CU3"))
    say(errs)
    assert not bad, bad
    # mismatched resolution is refused
    other = gpu_mod.GpuChannel(case.channel_case(96, 129, 192))
    with pytest.raises(ValueError):
        restart.load_flow(path, other)

"""The time-loop driver's host logic (CPU) and a short run on the GPU."""
import os

import numpy as np
import pytest

from diablo_amd import case, driver


def test_interval_due_steps_and_time():
    # positive: every N steps (for/diablo.f:93, 104)
    assert [driver.interval_due(3, ts, 0.0, 0.1) for ts in range(1, 7)] == [False, False, True, False, False, True]
    # negative: when TIME crosses a multiple of |N| within the next DELTA_T
    assert driver.interval_due(-2, 5, 1.95, 0.1)
    assert not driver.interval_due(-2, 5, 1.5, 0.1)
    assert driver.interval_due(-2, 5, 3.99, 0.02)


def test_end_run(tmp_path):
    d = str(tmp_path)
    assert driver.end_run(100.0, 50.0, d, now=120.0) is None
    assert driver.end_run(100.0, 10.0, d, now=120.0) == "wall-time hit"
    open(os.path.join(d, "stop.now"), "w").close()
    assert driver.end_run(100.0, 50.0, d, now=120.0) == "stop.now file"


def test_decks_round_trip(tmp_path):
    cfg = case.kh_case(32, 33, 24, kick=0.05).replace(
        variable_dt=True, cfl=0.4, update_dt=3, create_new_flow=False, time_limit=3600.0, reset_time=True,
        save_flow_int=7, save_stats_int=-2, n_steps=11, les=True)
    case.write_decks(cfg, str(tmp_path))
    dims = case.read_grid_def(str(tmp_path))
    assert dims == (32, 33, 24, 1)
    back = case.read_decks(str(tmp_path), *dims)
    for k in ("nu", "lx", "lz", "delta_t", "variable_dt", "cfl", "update_dt", "create_new_flow", "time_limit",
              "reset_time", "save_flow_int", "save_stats_int", "n_steps", "ri", "pr", "th_bc_ymin_c1", "f_type",
              "les"):
        assert getattr(back, k) == getattr(cfg, k), k
    # only the constant Smagorinsky model (LES_MODEL_TYPE = 1) is accepted
    path = os.path.join(str(tmp_path), "input_chan.dat")
    text = open(path).read().replace("C LES:   LES_MODEL_TYPE\n    1", "C LES:   LES_MODEL_TYPE\n    2")
    open(path, "w").write(text)
    with pytest.raises(ValueError):
        case.read_decks(str(tmp_path), *dims)


@pytest.mark.gpu
def test_driver_run(tmp_path):
    """Four steps with VARIABLE_DT, stats every 2 steps and a restart at step 3:
    the driver's state equals a hand-stepped context's, the text files have
    the reference's layout, and stop.now ends a run after one step."""
    from diablo_amd import gpu
    d = str(tmp_path)
    cfg = case.channel_case(32, 17, 24, kick=0.3, streak=0.1).replace(
        variable_dt=True, cfl=0.3, update_dt=1, save_stats_int=2, save_flow_int=3, n_steps=4)
    case.write_decks(cfg, d)
    r = driver.ChannelRun(d, log=lambda *a: None)
    r.initialize()
    out = r.run()
    assert out["steps"] == 4 and out["stopped"] is None
    ref = gpu.GpuChannel(r.cfg, r.gy)
    ref.initialize(*case.initial_physical(r.cfg, r.gy))
    ref.step(4, courant_cfl=0.3)
    assert ref.delta_t == r.chan.delta_t and ref.time == r.chan.time
    a, b = r.chan.state(), ref.state()
    for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
        assert np.array_equal(a[k], b[k]), k
    assert os.path.exists(os.path.join(d, "out.npz")) and os.path.exists(os.path.join(d, "end.npz"))
    lines = open(os.path.join(d, "mean.txt")).read().splitlines()
    # the new flow's record (CREATE_FLOW_CHAN's SAVE_STATS_CHAN, step 0),
    # steps 2 and 4, final (step 4): 4 blocks of 2 header lines + NY rows
    assert len(lines) == 4 * (2 + cfg.ny)
    steps = [int(lines[b * (2 + cfg.ny)].split()[0]) for b in range(4)]
    assert steps == [0, 2, 4, 4]
    row = lines[2].split()
    assert int(row[0]) == 1 and len(row) == 18
    tke = open(os.path.join(d, "tke.txt")).read().splitlines()
    assert len(tke) == 4 * (1 + cfg.ny - 2)
    open(os.path.join(d, "stop.now"), "w").close()
    r2 = driver.ChannelRun(d, log=lambda *a: None)
    r2.initialize()
    out2 = r2.run()
    # END_RUN stops after step 1; the final records carry TIME_STEP - 1
    # (for/diablo.f:130 after the loop's EXIT)
    assert out2["steps"] == 1 and out2["time_step"] == 0


@pytest.mark.gpu
def test_driver_les_spinup(tmp_path):
    """USE_LES on a new flow: NU_T stays zero for 100 steps, then les_chan
    runs in every substep (for/channel.f:303); the driver's state equals a
    context stepped by hand with the same switches."""
    from diablo_amd import gpu
    d = str(tmp_path)
    cfg = case.channel_case(32, 17, 24, kick=0.3, streak=0.1).replace(les=True, n_steps=102)
    case.write_decks(cfg, d)
    r = driver.ChannelRun(d, log=lambda *a: None)
    r.initialize()
    assert r.run()["steps"] == 102
    ref = gpu.GpuChannel(r.cfg, r.gy)
    ref.initialize(*case.initial_physical(r.cfg, r.gy))
    ref.set_les(True, model=False)
    ref.step(100)
    ref.set_les(True, model=True)
    ref.step(2)
    a, b = r.chan.state(), ref.state()
    for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
        assert np.array_equal(a[k], b[k]), k
    # tke_les.txt: the initial block (NU_T = 0) and the final one
    lines = open(os.path.join(d, "tke_les.txt")).read().splitlines()
    assert len(lines) == 2 * (1 + cfg.ny)
    first = [float(x.split()[2]) for x in lines[2:1 + cfg.ny]]
    assert all(v == 0.0 for v in first)
    last = np.array([float(x.split()[2]) for x in lines[cfg.ny + 3:]])
    assert np.all(np.isfinite(last)) and np.abs(last).max() > 0.0

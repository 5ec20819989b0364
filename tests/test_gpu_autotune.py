"""The launch-shape autotuner (dgpu_autotune): it leaves the state as it
found it, picks one of its candidates, and the planes-per-block choice does
not change the numbers."""
import numpy as np
import pytest

from diablo_amd import case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu_mod():
    from diablo_amd import gpu
    return gpu


def test_autotune_fp64_keeps_state_and_numbers(gpu_mod):
    cfg = case.channel_case(96, 65, 96, kick=0.3, streak=0.1)
    g = gpu_mod.GpuChannel(cfg)
    g.initialize(*case.initial_physical(cfg))
    g.step(2)
    st = g.state()
    r = g.autotune(2)
    say("fp64 autotune", r)
    assert r["xchunk"] in (4, 8, 16, 32, 64) and not r["zdft"] and r["ms_per_step"] > 0
    after = g.state()
    for k in st:
        assert np.array_equal(after[k], st[k]), k
    h = gpu_mod.GpuChannel(cfg)
    h.load_state(st)
    h.time = g.time
    g.step(1)
    h.step(1)
    a, b = g.state(), h.state()
    for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
        assert np.array_equal(a[k], b[k]), k


def test_autotune_mixed_keeps_state(gpu_mod):
    cfg = case.kh_case(128, 65, 128, kick=0.05, dt=2e-3)
    g = gpu_mod.GpuChannel(cfg, precision="mixed")
    g.initialize(*case.initial_physical(cfg))
    g.step(2)
    st = g.state()
    r = g.autotune(2)
    say("mixed autotune", r)
    assert r["xchunk"] in (4, 8, 16, 32, 64)
    after = g.state()
    for k in st:
        assert np.array_equal(after[k], st[k]), k
    g.step(2)
    assert np.all(np.isfinite(g.state()["CU1"]))

"""The launch-shape autotuner (dgpu_autotune): it leaves the state as it
found it, picks one of its candidates, and the planes-per-block choice does
not change the numbers (a context with the same column kernels gives the
same bits)."""
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
    # the same column kernels as the tuned context (FP64 also tries the
    # fused column kernel, which rounds differently from the two kernels)
    h = gpu_mod.GpuChannel(cfg)
    h.set_column_shape(r["column_waves"], 1 if r["column_fused"] else 0)
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


def test_ensemble_autotune_keeps_states(gpu_mod):
    """dgpu_ensemble_autotune leaves every member's state as it was, picks a
    shape from its candidates, and the ensemble then still advances each
    member bit for bit as its single run (with that member's shape)."""
    cfgs = [case.kh_case(24, 17, 24, kick=0.05, dt=0.01, ri=0.1, seed=1),
            case.kh_case(48, 33, 36, kick=0.05, dt=0.01, ri=0.2, re=500.0, seed=2),
            case.kh_case(48, 65, 36, kick=0.05, dt=0.01, ri=0.15, seed=3)]
    for precision in ("fp64", "mixed"):
        ens = gpu_mod.GpuEnsemble(cfgs, precision=precision)
        singles = []
        for cfg, m in zip(cfgs, ens.members):
            phys = case.initial_physical(cfg)
            m.initialize(*phys)
            s = gpu_mod.GpuChannel(cfg, precision=precision)
            s.initialize(*phys)
            singles.append(s)
        ens.step(1)
        ens.synchronize()
        before = [m.state() for m in ens.members]
        r = ens.autotune(1)
        say(precision, "ensemble autotune", r)
        assert r["xchunk"] in (8, 16, 32, 64) and r["column_waves"] in (4, 8, 16) and r["ms_per_step"] > 0
        for b, m in zip(before, ens.members):
            a = m.state()
            for k in ("CU1", "CU2", "CU3", "CP", "CTH0", "CF1", "CF2", "CF3", "CFTH0"):
                assert np.array_equal(a[k], b[k]), (precision, k)
        for s, m in zip(singles, ens.members):
            s.step(1)
            s.set_column_shape(m.column_shape(), -1)
        ens.step(2)
        ens.synchronize()
        for s in singles:
            s.step(2)
            s.synchronize()
        for cfg, s, m in zip(cfgs, singles, ens.members):
            a, b = m.state(), s.state()
            for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
                assert np.array_equal(a[k], b[k]), (precision, cfg.nx, cfg.ny, k)
        ens.close()
        for s in singles:
            s.close()

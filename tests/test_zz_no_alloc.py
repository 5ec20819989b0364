"""dgpu_stats_chan_noalloc, the entry point that allocates nothing, against
dgpu_stats_chan (same kernel; the per-plane sums go through the library's
fixed device array instead of a buffer allocated on the first call).  Both
are called through the C-ABI; the results must be bit-identical, on
consecutive calls whose second input is smaller than the first (the fixed
array keeps the previous call's contents)."""
import numpy as np
import pytest

from diablo_amd import case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _stepped(gpu, cfg, precision="fp64", steps=2):
    g = gpu.GpuChannel(cfg, precision=precision)
    g.initialize(*case.initial_physical(cfg))
    g.step(steps)
    g.synchronize()
    return g


def _same(a, ua, b, ub):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert ua == ub


def test_noalloc_stats_match_allocating_entry_point():
    from diablo_amd import gpu
    cases = [
        (case.channel_case(48, 65, 36, kick=0.05, streak=0.05), "fp64"),
        (case.kh_case(32, 33, 24, kick=0.05), "fp64"),
        (case.channel_case(24, 17, 18, kick=0.05), "mixed"),
    ]
    for n, (cfg, prec) in enumerate(cases):
        say(f"noalloc stats: call {n + 1}, {cfg.nx}x{cfg.ny}x{cfg.nz} {prec}")
        g = _stepped(gpu, cfg, prec)
        if n % 2 == 0:
            b, ub = g.stats_chan_noalloc()
            a, ua = g.stats_chan()
        else:
            a, ua = g.stats_chan()
            b, ub = g.stats_chan_noalloc()
        _same(a, ua, b, ub)
        assert np.all(np.isfinite(b["urms"])) and b["urms"].max() > 0
        g.close()


def test_noalloc_stats_repeat_on_one_context():
    from diablo_amd import gpu
    cfg = case.kh_case(32, 33, 24, kick=0.05)
    g = _stepped(gpu, cfg)
    first = g.stats_chan_noalloc()
    g.step(1)
    g.synchronize()
    second = g.stats_chan_noalloc()
    ref = g.stats_chan()
    _same(ref[0], ref[1], second[0], second[1])
    assert not np.array_equal(first[0]["urms"], second[0]["urms"])
    g.close()


def test_noalloc_stats_refuses_tall_grid():
    from diablo_amd import gpu
    cfg = case.channel_case(16, 4097, 16, kick=0.0)
    say("noalloc stats: NY = 4097 must be refused")
    g = gpu.GpuChannel(cfg)
    with pytest.raises(gpu.DgpuError, match="DGPU_NOALLOC_MAX_NY"):
        g.stats_chan_noalloc()
    g.close()

"""The mixed path's z transforms as split-BF16 DFT-as-GEMM on the matrix
cores (dgpu_set_zdft): against the same path with the FFT z transforms after
one step, against the FP64 path after 20 steps, replayed graphs against eager
launches, and the refused configurations."""
import numpy as np
import pytest

from diablo_amd import case
from gpu_util import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("CU1", "CU2", "CU3", "CP", "CTH0")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu_mod():
    from diablo_amd import gpu
    return gpu


def _rel_l2(cfg, a, b):
    ka, kb = case.kept(cfg, a), case.kept(cfg, b)
    return float(np.linalg.norm(ka - kb) / np.linalg.norm(kb))


@pytest.mark.parametrize("dims", [(64, 33, 64), (96, 33, 96), (128, 65, 128), (256, 65, 256)])
def test_zdft_step_matches_fft(gpu_mod, dims):
    """One step from the same state: split-BF16 products keep about 16
    significant bits, so the velocity and scalar agree with the FFT
    transforms of the same path within 1e-5 (relative L2 over the kept
    modes)."""
    cfg = case.kh_case(*dims, kick=0.05, dt=2e-3)
    say("zdft vs fft", dims)
    a = gpu_mod.GpuChannel(cfg, precision="mixed")
    a.initialize(*case.initial_physical(cfg))
    a.step(2)
    st = a.state()
    b = gpu_mod.GpuChannel(cfg, precision="mixed")
    b.load_state(st)
    b.time = a.time
    b.set_zdft(True)
    a.step(1)
    b.step(1)
    sa, sb = a.state(), b.state()
    errs = {k: _rel_l2(cfg, sb[k], sa[k]) for k in FIELDS}
    say(errs)
    for k in FIELDS:
        if k != "CP":
            assert errs[k] < 1e-5, (k, errs[k])
    # CP = CP + phi / H_BAR: phi is the small projection potential, whose
    # relative round-off the division carries into CP undamped
    assert errs["CP"] < 1e-2, errs["CP"]
    # the step moved the state: the comparison is not of unchanged fields
    assert _rel_l2(cfg, sa["CU2"], st["CU2"]) > 1e-4


def test_zdft_against_fp64(gpu_mod):
    """20 steps of the KH layer at 128x65x128: relative L2 of the velocity
    against the FP64 path stays within 1e-5."""
    cfg = case.kh_case(128, 65, 128, kick=1e-3, dt=2e-3)
    say("zdft vs fp64")
    phys = case.initial_physical(cfg)
    g64 = gpu_mod.GpuChannel(cfg)
    g64.initialize(*phys)
    gz = gpu_mod.GpuChannel(cfg, precision="mixed")
    gz.initialize(*phys)
    gz.set_zdft(True)
    g64.step(20)
    gz.step(20)
    s64, sz = g64.state(), gz.state()
    num = sum(np.linalg.norm(case.kept(cfg, sz[k]) - case.kept(cfg, s64[k])) ** 2 for k in ("CU1", "CU2", "CU3"))
    den = sum(np.linalg.norm(case.kept(cfg, s64[k])) ** 2 for k in ("CU1", "CU2", "CU3"))
    e = float(np.sqrt(num / den))
    say("rel-L2 velocity", e)
    assert np.isfinite(e) and e < 1e-5, e


def test_zdft_without_scalar(gpu_mod):
    """N_TH = 0: three fields through the GEMM z transforms."""
    cfg = case.channel_case(96, 33, 96, kick=0.3, streak=0.1, n_th=0)
    say("zdft n_th=0")
    a = gpu_mod.GpuChannel(cfg, precision="mixed")
    a.initialize(*case.initial_physical(cfg))
    a.step(1)
    st = a.state()
    b = gpu_mod.GpuChannel(cfg, precision="mixed")
    b.load_state(st)
    b.time = a.time
    b.set_zdft(True)
    a.step(1)
    b.step(1)
    sa, sb = a.state(), b.state()
    for k in ("CU1", "CU2", "CU3"):
        e = _rel_l2(cfg, sb[k], sa[k])
        say(k, e)
        assert e < 1e-5, (k, e)


def test_zdft_graph_replays_match_eager(gpu_mod):
    cfg = case.kh_case(64, 33, 64, kick=0.05, dt=2e-3)
    say("zdft graphs")
    phys = case.initial_physical(cfg)
    a = gpu_mod.GpuChannel(cfg, precision="mixed")
    a.initialize(*phys)
    a.set_zdft(True)
    b = gpu_mod.GpuChannel(cfg, precision="mixed")
    b.initialize(*phys)
    b.set_zdft(True)
    b.set_graphs(False)
    for n in range(3):
        a.step(1)
        b.step(1)
        sa, sb = a.state(), b.state()
        for k in FIELDS:
            assert np.array_equal(sa[k], sb[k]), (n, k)


def test_zdft_refused(gpu_mod):
    cfg = case.kh_case(64, 33, 64)
    with pytest.raises(RuntimeError):
        gpu_mod.GpuChannel(cfg).set_zdft(True)             # FP64 context
    with pytest.raises(RuntimeError):
        gpu_mod.GpuChannel(case.kh_case(64, 33, 384), precision="mixed").set_zdft(True)   # NZ > 256
    g = gpu_mod.GpuChannel(cfg, precision="mixed")
    g.set_zdft(True)
    g.set_zdft(False)                                      # back to the FFTs

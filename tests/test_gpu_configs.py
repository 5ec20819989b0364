"""Parity with the CPU oracle at the workloads' own sizes and on the batched
ensemble's member grids, through the C-ABI:

* the 384 x 257 x 384 channel (Re 395, streak seed) and the 256 x 193 x 256
  Kelvin-Helmholtz layer, one RK3 step from the same seeded state, FP64 at
  1e-10 (these sizes run the compile-time transform plans of 384 and 256),
  and the mixed path's velocity within 1e-6 relative L2 after that step;
* ensemble members on NX = NZ in {64, 96, 128, 192} (the batched launches'
  compile-time plans of those lengths), FP64 at 1e-10 and mixed within 1e-6
  after one step, each also against its own single run;
* the mixed paths against the oracle directly after 20 steps (the 1e-5
  relative L2 bar).

The oracle runs on the host's cores (OpenMP)."""
import pytest

import numpy as np

from diablo_amd import case, metrics
from gpu_util import compare_states, rel_err

pytestmark = pytest.mark.gpu

FP64_FIELDS = ("CU1", "CU2", "CU3", "CP", "CTH0", "CF1", "CF2", "CF3", "CFTH0")


def say(*a):
    print(*a, flush=True)


def _oracle(oracle_mod, cfg, phys, steps):
    o = oracle_mod.Oracle(cfg)
    o.initialize(*phys)
    for _ in range(steps):
        o.step()
    return o


def _vel_rel_l2(g_ref, s_ref, g, s):
    return metrics.rel_l2(metrics.physical_velocity(g, s), metrics.physical_velocity(g_ref, s_ref), g.cfg.ny)


@pytest.mark.parametrize("workload", ["channel384", "kh256"])
def test_workload_step_vs_oracle(oracle_mod, workload):
    """The bar of each field is 1e-10 of its max, or ten times the oracle's
    own spread when its start state moves by 1e-16 (relative), whichever is
    larger: a field much smaller than the products it is made of (the
    channel's CFTH, 4e-5 against O(1) products of TH and U) carries their
    round-off (3e-10 of its max in the oracle itself at 384)."""
    from diablo_amd import gpu
    cfg = case.WORKLOADS[workload]()
    say(f"{workload}: {cfg.nx}x{cfg.ny}x{cfg.nz}, initial state")
    phys = case.initial_physical(cfg)
    say(f"{workload}: oracle, one step")
    o = oracle_mod.Oracle(cfg)
    o.initialize(*phys)
    s0 = o.state()
    o.step()
    so = o.state()
    say(f"{workload}: oracle from the perturbed state, one step")
    rng = np.random.default_rng(99)
    o.load_state({k: v * (1 + 1e-16 * rng.standard_normal(v.shape)) for k, v in s0.items()})
    o.time = 0.0
    o.step()
    sp = o.state()
    del o
    sens = {k: rel_err(cfg, sp[k], so[k]) for k in FP64_FIELDS}
    say(f"{workload}: FP64 GPU, one step")
    g = gpu.GpuChannel(cfg)
    g.initialize(*phys)
    g.step(1)
    g.synchronize()
    sg = g.state()
    errs = {k: rel_err(cfg, sg[k], so[k]) for k in FP64_FIELDS}
    say("errors", errs)
    say("oracle self-sensitivity", sens)
    bad = {k: e for k, e in errs.items() if not e <= max(1e-10, 10 * sens[k])}
    assert not bad, bad
    say(f"{workload}: mixed GPU, one step")
    gm = gpu.GpuChannel(cfg, precision="mixed")
    gm.initialize(*phys)
    gm.step(1)
    gm.synchronize()
    e = _vel_rel_l2(g, so, gm, gm.state())
    say(f"{workload}: mixed velocity rel-L2 vs oracle {e:.3e}")
    assert e < 1e-6
    gm.close()
    g.close()


def _member_grids():
    cfgs = []
    for m, n in enumerate((64, 96, 128, 192)):
        for ny in (17, 33):
            cfgs.append(case.kh_case(n, ny, n, re=(500.0, 1000.0)[m % 2], ri=0.05 + 0.05 * m, dt=2e-3,
                                     kick=0.05, seed=40 + len(cfgs)))
    return cfgs


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_member_grids_vs_oracle(oracle_mod, precision):
    from diablo_amd import gpu
    cfgs = _member_grids()
    phys = [case.initial_physical(c) for c in cfgs]
    say(f"member grids {precision}: {len(cfgs)} members")
    ens = gpu.GpuEnsemble(cfgs, precision=precision)
    for p, m in zip(phys, ens.members):
        m.initialize(*p)
    ens.step(1)
    ens.synchronize()
    for cfg, p, m in zip(cfgs, phys, ens.members):
        say(f"  member {cfg.nx}x{cfg.ny}x{cfg.nz}")
        so = _oracle(oracle_mod, cfg, p, 1).state()
        single = gpu.GpuChannel(cfg, precision=precision)
        single.initialize(*p)
        single.step(1)
        single.synchronize()
        sm, ss = m.state(), single.state()
        if precision == "fp64":
            errs, bad = compare_states(cfg, sm, so, 1e-10, fields=FP64_FIELDS)
            assert not bad, (cfg.nx, cfg.ny, bad)
            errs, bad = compare_states(cfg, sm, ss, 1e-12, fields=FP64_FIELDS)
            assert not bad, (cfg.nx, cfg.ny, "vs single run", bad)
        else:
            g64 = gpu.GpuChannel(cfg)
            e = _vel_rel_l2(g64, so, m, sm)
            es = _vel_rel_l2(g64, ss, m, sm)
            say(f"    mixed rel-L2 vs oracle {e:.2e}, vs single mixed run {es:.2e}")
            assert e < 1e-6 and es < 1e-6
            g64.close()
        single.close()
    ens.close()


@pytest.mark.parametrize("precision", ["mixed", "mixed_f64state"])
def test_mixed_vs_oracle_20_steps(oracle_mod, precision):
    from diablo_amd import gpu
    cfg = case.channel_case(48, 33, 48, re=180.0, dt=1e-3, kick=0.05, streak=0.05)
    phys = case.initial_physical(cfg)
    so = _oracle(oracle_mod, cfg, phys, 20).state()
    g64 = gpu.GpuChannel(cfg)
    gm = gpu.GpuChannel(cfg, precision=precision)
    gm.initialize(*phys)
    gm.step(20)
    gm.synchronize()
    r = metrics.compare(g64, so, gm, gm.state())
    say(f"{precision} after 20 steps vs oracle: {r}")
    assert r["rel_l2"] < 1e-5 and r["ke_drift"] < 1e-5
    assert r["max_div"] < 1e-6
    gm.close()
    g64.close()

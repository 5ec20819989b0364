"""Ensemble stepping (dgpu_ensemble_rk_chan_1) under a changing time step and
with a member's COURANT request, against each member's single run, through
the C-ABI.  Members with the same NX, NZ share batched launches; every
member must advance bit for bit as the same run alone."""
import numpy as np
import pytest

from diablo_amd import case

pytestmark = pytest.mark.gpu

FIELDS = ("CU1", "CU2", "CU3", "CP", "CTH0")


def say(*a):
    print(*a, flush=True)


def _cfgs():
    return [case.kh_case(24, 17, 24, kick=0.05, dt=0.01, ri=0.1, seed=1),
            case.kh_case(24, 33, 24, kick=0.05, dt=0.01, ri=0.15, seed=4),
            case.kh_case(48, 33, 36, kick=0.05, dt=0.01, ri=0.2, re=500.0, seed=2),
            case.kh_case(48, 17, 36, kick=0.05, dt=0.005, ri=0.05, seed=5)]


def _pair(gpu, cfgs, precision):
    ens = gpu.GpuEnsemble(cfgs, precision=precision)
    singles = []
    for cfg, m in zip(cfgs, ens.members):
        phys = case.initial_physical(cfg)
        m.initialize(*phys)
        g = gpu.GpuChannel(cfg, precision=precision)
        g.set_graphs(False)
        g.initialize(*phys)
        singles.append(g)
    ens.synchronize()
    return ens, singles


def _same(cfgs, ens, singles, tag):
    for cfg, m, g in zip(cfgs, ens.members, singles):
        a, b = m.state(), g.state()
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), (tag, cfg.nx, cfg.ny, k)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_new_dt_each_step_without_sync(precision):
    """A new H_BAR on every step, enqueued back to back with no
    synchronisation between the ensemble's steps: each stage's
    Crank-Nicolson tables are rebuilt only after the ensemble's earlier
    launches that read them (graphs and batched launches on the ensemble's
    stream) have finished."""
    from diablo_amd import case as cs, gpu
    cfgs = _cfgs()
    ens, singles = _pair(gpu, cfgs, precision)
    dts = [1.0, 0.7, 1.3, 0.9, 1.1]
    for s, f in enumerate(dts):
        say(f"ensemble {precision}: step {s} with dt x {f}")
        ens.h_bar = np.ascontiguousarray(np.concatenate([cs.rkw3_table(c.delta_t * f)[0] for c in cfgs]))
        ens.step(1)
        for cfg, g in zip(cfgs, singles):
            g.set_delta_t(cfg.delta_t * f)
            g.step(1)
    ens.synchronize()
    for g in singles:
        g.synchronize()
    _same(cfgs, ens, singles, "after the dt sequence")
    ens.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_member_courant(precision):
    """dgpu_courant_request on one member before the ensemble's RK stage 3:
    the maxima are recorded (dgpu_courant succeeds), DELTA_T equals the
    single run's, and the states stay bit-identical on the following step
    (graphs and batched launches again)."""
    from diablo_amd import gpu
    cfgs = _cfgs()
    ens, singles = _pair(gpu, cfgs, precision)
    cfl = 0.5
    ens.step(1)
    for g in singles:
        g.step(1)
    for rep in range(2):
        say(f"ensemble {precision}: COURANT on member 1, step {rep}")
        ens.rk_chan_1(1)
        ens.rk_chan_1(2)
        m = ens.members[1]
        m._chk(m.L.dgpu_courant_request(m.h), "dgpu_courant_request")
        ens.rk_chan_1(3)
        ens.synchronize()
        dt_e = m.courant(cfl)
        g = singles[1]
        g.step(1, courant_cfl=cfl)
        for i, gg in enumerate(singles):
            if i != 1:
                gg.step(1)
        assert dt_e == g.delta_t, (dt_e, g.delta_t)
        ens.h_bar[3:6] = m.h_bar
        _same(cfgs, ens, singles, f"COURANT step {rep}")
    ens.step(1)
    for g in singles:
        g.step(1)
    ens.synchronize()
    _same(cfgs, ens, singles, "step after COURANT")
    ens.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_ensemble_group_streams(monkeypatch, graphs):
    """The member groups (here two: NX = 24 and 48) run on streams of their
    own, forked from and joined into the ensemble stream; three steps back
    to back with no synchronisation between them, eagerly and through the
    stage graphs, give each member's single run bit for bit, and the same
    bits as the ensemble on one stream (DGPU_ENS_STREAMS=0)."""
    from diablo_amd import gpu
    cfgs = _cfgs()
    runs = []
    for one in (False, True):
        if one:
            monkeypatch.setenv("DGPU_ENS_STREAMS", "0")
        else:
            monkeypatch.delenv("DGPU_ENS_STREAMS", raising=False)
        ens, singles = _pair(gpu, cfgs, "mixed")
        ens.set_graphs(graphs)
        say(f"ensemble, graphs {graphs}, one stream {one}")
        ens.step(3)
        ens.synchronize()
        for g in singles:
            g.step(3)
            g.synchronize()
        _same(cfgs, ens, singles, f"graphs {graphs} one stream {one}")
        runs.append([m.state() for m in ens.members])
        ens.close()
        for g in singles:
            g.close()
    for a, b in zip(*runs):
        for k in FIELDS:
            assert np.array_equal(a[k], b[k]), k

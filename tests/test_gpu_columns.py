"""The column stage's launch shapes (dgpu_set_column_shape): the segmented
column solves in blocks of 4, 8 and 16 waves give the same bits; the
whole-column sweeps agree to round-off; the fused one-kernel column stage
agrees with the two kernels to round-off in FP64 and is as accurate as they
are with the FP32 state, and agrees with the oracle; the oracle agrees with
the shape the context runs; shapes a column does not fit are refused.  All
through the C-ABI."""
import numpy as np
import pytest

from diablo_amd import case
from gpu_util import compare_states, kitchen_sink, rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("CU1", "CU2", "CU3", "CP", "CF1", "CF2", "CF3", "CTH0", "CFTH0")


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu_mod():
    from diablo_amd import gpu
    return gpu


def _run(gpu, cfg, precision, waves, steps=3, fused=0):
    g = gpu.GpuChannel(cfg, precision=precision)
    g.set_column_shape(waves, fused)
    assert g.column_shape() == waves
    assert g.column_fused() == (fused == 1)
    g.initialize(*case.initial_physical(cfg))
    g.step(steps)
    g.synchronize()
    s = g.state()
    g.close()
    return s


CASES = {
    "channel": lambda: case.channel_case(48, 65, 36, re=395.0, dt=5e-4, kick=0.05, streak=0.1),
    "kh": lambda: case.kh_case(32, 33, 24, kick=0.05, dt=0.01, ri=0.1),
    "kitchen": lambda: kitchen_sink(),
}


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_f64state"])
@pytest.mark.parametrize("name", list(CASES))
def test_column_shapes_same_bits(gpu_mod, name, precision):
    """4, 8 and 16 waves per block: the same bits after 3 steps (each mode
    goes through the same operations in the same order).  The whole-column
    sweeps: within 1e-11 of them with the FP64 column solves (the parity bar
    of the substep tests; CP, the sum of the projections over the run, is
    the most sensitive at ~2e-12).  With the FP32 state both column
    variants are compared with the FP64 run instead: the segmented solves
    must be no less accurate than the sweeps (within 1.5x of the sweeps'
    error, field by field, plus 1e-7)."""
    cfg = CASES[name]()
    runs = {}
    for w in (16, 8, 4, 0):
        say(f"{name} {precision}: column shape {w} waves")
        runs[w] = _run(gpu_mod, cfg, precision, w)
    for w in (8, 4):
        for k in FIELDS:
            assert np.array_equal(runs[w][k], runs[16][k]), (w, k)
    if precision != "mixed":
        errs, bad = compare_states(cfg, runs[0], runs[16], 1e-11, FIELDS)
        say("sweeps vs segmented", errs)
        assert not bad, bad
        return
    ref = _run(gpu_mod, cfg, "fp64", 16)
    e_seg, _ = compare_states(cfg, runs[16], ref, 1.0, FIELDS)
    e_swp, _ = compare_states(cfg, runs[0], ref, 1.0, FIELDS)
    say("segmented vs FP64", e_seg)
    say("sweeps vs FP64", e_swp)
    bad = {k: (e_seg[k], e_swp[k]) for k in e_seg if not e_seg[k] <= 1.5 * e_swp[k] + 1e-7}
    assert not bad, bad


@pytest.mark.parametrize("waves,fused", [(4, 0), (8, 0), (8, 1)])
def test_column_shape_vs_oracle(gpu_mod, oracle_mod, waves, fused):
    """The kitchen-sink deck (every term of RK_CHAN_1 on) through the
    smaller blocks and the FP64 fused column kernel against the oracle,
    1e-11 after 2 steps."""
    cfg = kitchen_sink()
    phys = case.initial_physical(cfg)
    g = gpu_mod.GpuChannel(cfg)
    g.set_column_shape(waves, fused)
    assert g.column_fused() == (fused == 1)
    g.initialize(*phys)
    g.step(2)
    o = oracle_mod.Oracle(cfg)
    o.initialize(*phys)
    for _ in range(2):
        o.step()
    errs, bad = compare_states(cfg, g.state(), o.state(), 1e-11)
    g.close()
    say(errs)
    assert not bad, bad


def test_column_shape_refusals(gpu_mod):
    cfg = case.channel_case(24, 17, 18, kick=0.05)
    g = gpu_mod.GpuChannel(cfg)
    for bad in (1, 2, 5, 32, -1):
        with pytest.raises(gpu_mod.DgpuError, match="waves per block"):
            g.set_column_shape(bad)
    g.close()
    # 1025 rows: 33 rows per FP64 segment, beyond the segmented kernels
    tall = case.channel_case(24, 1025, 18, kick=0.0, streak=0.0)
    t = gpu_mod.GpuChannel(tall)
    assert t.column_shape() == 0
    with pytest.raises(gpu_mod.DgpuError, match="do not fit"):
        t.set_column_shape(8)
    t.set_column_shape(0)
    t.close()
    # the fused kernel: the FP32 state or the FP64 path (not the FP64 state
    # under FP32 transforms), NY <= 268
    f = gpu_mod.GpuChannel(cfg, precision="mixed_f64state")
    with pytest.raises(gpu_mod.DgpuError, match="FP32 state"):
        f.set_column_shape(8, 1)
    f.close()
    m = gpu_mod.GpuChannel(case.channel_case(24, 289, 18, kick=0.0, streak=0.0), precision="mixed")
    assert not m.column_fused()
    with pytest.raises(gpu_mod.DgpuError, match="column shape"):
        m.set_column_shape(8, 1)
    m.close()
    # by default from NY = 160 up, and with the FP32 state on the short
    # columns it cuts into 16 segments (NY <= 142)
    for ny, want in ((129, True), (150, False), (161, True)):
        a = gpu_mod.GpuChannel(case.channel_case(24, ny, 18, kick=0.0, streak=0.0), precision="mixed")
        assert a.column_fused() == want, ny
        a.close()
    # FP64: fused by default from NY = 160 at 8 waves only; not at 16 waves
    # (its registers do not fit a block of 16)
    d = gpu_mod.GpuChannel(case.channel_case(24, 161, 18, kick=0.0, streak=0.0))
    assert d.column_shape() == 8 and d.column_fused()
    d.set_column_shape(4)
    assert not d.column_fused()
    d.set_column_shape(4, 1)
    assert d.column_fused()
    with pytest.raises(gpu_mod.DgpuError, match="column shape"):
        d.set_column_shape(16, 1)
    d.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("name", list(CASES))
def test_fused_column_stage(gpu_mod, name, precision):
    """The fused column kernel: 4, 8 and (FP32 state) 16 waves give the same
    bits (30 segments at every shape).  In FP64 it agrees with the two
    kernels within 1e-11 (the parity bar of the substep tests; the two
    kernels solve on 32 segments).  With the FP32 state, against the FP64 run it is no
    less accurate than the two kernels (within 1.5x of their error, field by
    field, plus 1e-7)."""
    cfg = CASES[name]()
    runs = {}
    shapes = (8, 4, 16) if precision == "mixed" else (8, 4)
    for w in shapes:
        say(f"{name} {precision}: fused column stage, {w} waves")
        runs[w] = _run(gpu_mod, cfg, precision, w, fused=1)
    for w in shapes[1:]:
        for k in FIELDS:
            assert np.array_equal(runs[w][k], runs[8][k]), (w, k)
    two = _run(gpu_mod, cfg, precision, 8, fused=0)
    if precision == "fp64":
        errs, bad = compare_states(cfg, runs[8], two, 1e-11, FIELDS)
        say("fused vs two kernels", errs)
        assert not bad, bad
        return
    ref = _run(gpu_mod, cfg, "fp64", 8)
    e_f, _ = compare_states(cfg, runs[8], ref, 1.0, FIELDS)
    e_t, _ = compare_states(cfg, two, ref, 1.0, FIELDS)
    say("fused vs FP64", e_f)
    say("two kernels vs FP64", e_t)
    bad = {k: (e_f[k], e_t[k]) for k in e_f if not e_f[k] <= 1.5 * e_t[k] + 1e-7}
    assert not bad, bad


def test_fused_column_stage_vs_oracle(gpu_mod, oracle_mod):
    """The fused column kernel against the oracle over 20 steps of the
    48 x 33 x 48 KH layer: rel-L2 of the velocity and scalar below 1e-5, as
    the mixed path's own bar."""
    from diablo_amd import metrics
    cfg = case.kh_case(48, 33, 48, kick=0.05, dt=0.01, ri=0.1)
    phys = case.initial_physical(cfg)
    g = gpu_mod.GpuChannel(cfg, precision="mixed")
    g.set_column_shape(8, 1)
    g.initialize(*phys)
    g.step(20)
    g.synchronize()
    o = oracle_mod.Oracle(cfg)
    o.initialize(*phys)
    for _ in range(20):
        o.step()
    sg, so = g.state(), o.state()
    g.close()
    for k in ("CU1", "CU2", "CU3", "CTH0"):
        a, b = case.kept(cfg, sg[k]), case.kept(cfg, so[k])
        rel = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
        say(k, rel)
        assert rel < 1e-5, (k, rel)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fused_column_stage_with_courant(gpu_mod, precision):
    """COURANT steps run the two kernels for their stage 3 (the maxima are
    taken between them); the run with the fused kernel follows the run with
    the two kernels: DELTA_T within 1e-6, the state within 1e-5."""
    cfg = case.kh_case(32, 33, 24, kick=0.05, dt=0.01, ri=0.1)
    phys = case.initial_physical(cfg)
    out = []
    for fz in (1, 0):
        g = gpu_mod.GpuChannel(cfg, precision=precision)
        g.set_column_shape(8, fz)
        g.initialize(*phys)
        dts = []
        for _ in range(4):
            g.step(1, courant_cfl=0.5)
            dts.append(g.delta_t)
        g.synchronize()
        out.append((dts, g.state()))
        g.close()
    (d1, s1), (d0, s0) = out
    say("dt fused", d1, "two kernels", d0)
    assert np.allclose(d1, d0, rtol=1e-6, atol=0)
    errs, bad = compare_states(cfg, s1, s0, 1e-5, ("CU1", "CU2", "CU3", "CTH0"))
    say(errs)
    assert not bad, bad


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ensemble_follows_member_shapes(gpu_mod, precision):
    """Ensemble graphs are captured per launch shape: after the members'
    column shape changes between steps, the ensemble re-captures and still
    matches each member's single run bit for bit (also through the fused
    column kernel, batched and single)."""
    cfgs = [case.kh_case(24, 17, 24, kick=0.05, dt=0.01, ri=0.1, seed=1),
            case.kh_case(48, 33, 36, kick=0.05, dt=0.01, ri=0.2, re=500.0, seed=2),
            case.kh_case(48, 65, 36, kick=0.05, dt=0.01, ri=0.15, seed=3)]
    ens = gpu_mod.GpuEnsemble(cfgs, precision=precision)
    singles = []
    for cfg, m in zip(cfgs, ens.members):
        phys = case.initial_physical(cfg)
        m.initialize(*phys)
        s = gpu_mod.GpuChannel(cfg, precision=precision)
        s.initialize(*phys)
        singles.append(s)
    ens.synchronize()
    # a step that changes only the fused setting must re-capture too
    shapes = ((16, 0), (8, 0), (8, 1), (4, 0)) if precision == "fp64" else ((16, 0), (8, 1), (4, 1), (4, 0))
    for step, (w, fz) in enumerate(shapes):
        say(f"ensemble {precision} step {step}: members at {w} waves, fused {fz}")
        for m, s in zip(ens.members, singles):
            m.set_column_shape(w, fz)
            s.set_column_shape(w, fz)
        ens.step(1)
        ens.synchronize()
        for s in singles:
            s.step(1)
    for cfg, m, s in zip(cfgs, ens.members, singles):
        s.synchronize()
        a, b = m.state(), s.state()
        for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
            assert np.array_equal(a[k], b[k]), (cfg.nx, k)
        assert rel_err(cfg, a["CU1"], b["CU1"]) == 0.0
    ens.close()
    for s in singles:
        s.close()

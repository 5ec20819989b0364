"""les_chan's constant Smagorinsky model on the device (dgpu_les_chan)
against the oracle: NU_T, the SGS force added to CF1-3 and the velocity BC
rows it sets, on a wall-bounded channel, a free-slip shear layer and the
deck with Neumann and moving-wall conditions; at 192x129x192 the model's
homogeneity (NU_T linear, the force quadratic in the velocity)."""
import numpy as np
import pytest

from diablo_amd import case
from gpu_util import FIELDS, compare_states, kitchen_sink, rel_err

pytestmark = pytest.mark.gpu

CASES = {
    "channel": lambda: case.channel_case(48, 33, 36, kick=0.3, streak=0.1),
    "kh": lambda: case.kh_case(32, 33, 24, kick=0.05),
    "sink": lambda: kitchen_sink(48, 25, 36),
    "no_scalar": lambda: case.channel_case(48, 33, 36, kick=0.3, streak=0.1, n_th=0),
}


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def gpu_mod():
    from diablo_amd import gpu
    return gpu


def _cf(cfg, st):
    return [case.kept(cfg, st[k]).copy() for k in ("CF1", "CF2", "CF3")]


@pytest.mark.parametrize("maker", list(CASES))
def test_les_chan_matches_oracle(oracle_mod, gpu_mod, maker):
    cfg = CASES[maker]()
    say("les_chan", maker, cfg.nx, cfg.ny, cfg.nz)
    o = oracle_mod.Oracle(cfg)
    o.initialize(*case.initial_physical(cfg))
    o.step()
    o.step()
    g = gpu_mod.GpuChannel(cfg)
    g.load_state(o.state())
    before = _cf(cfg, o.state())
    nut_o = o.les_chan()
    nut_g = g.les_chan()
    scale = np.abs(nut_o).max()
    assert scale > 0.0
    err = np.abs(nut_g - nut_o).max() / scale
    say("NU_T", err)
    assert err < 1e-12, err
    ny = cfg.ny
    for j in (0, 1, 2, ny, ny + 1):
        assert np.abs(nut_g[j]).max() == 0.0, j
    assert np.abs(nut_g[:, cfg.nz:, :]).max() == 0.0 and np.abs(nut_g[:, :, cfg.nx:]).max() == 0.0
    so, sg = o.state(), g.state()
    for q, (b, ao, ag) in enumerate(zip(before, _cf(cfg, so), _cf(cfg, sg))):
        d_o, d_g = ao - b, ag - b
        ref = np.abs(d_o).max()
        assert ref > 0.0, q
        e = np.abs(d_g - d_o).max() / ref
        say(f"dCF{q + 1}", e)
        assert e < 1e-11, (q, e)
    errs, bad = compare_states(cfg, sg, so, 1e-13, fields=("CU1", "CU2", "CU3"))
    say(errs)
    assert not bad, bad


def test_les_chan_mixed_context(oracle_mod, gpu_mod):
    """A mixed-precision context evaluates the model on its FP64 master state:
    the same result as an FP64 context loaded alike."""
    cfg = CASES["channel"]()
    o = oracle_mod.Oracle(cfg)
    o.initialize(*case.initial_physical(cfg))
    o.step()
    g = gpu_mod.GpuChannel(cfg)
    gm = gpu_mod.GpuChannel(cfg, precision="mixed")
    g.load_state(o.state())
    gm.load_state(o.state())
    say("les_chan mixed context")
    n64, nm = g.les_chan(), gm.les_chan()
    assert np.array_equal(n64, nm)
    a, b = g.state(), gm.state()
    for k in ("CF1", "CF2", "CF3", "CU1", "CU3"):
        assert np.array_equal(case.kept(cfg, a[k]), case.kept(cfg, b[k])), k


def test_les_chan_homogeneity_large(gpu_mod):
    """At 192x129x192 after two steps: scaling the velocity by 2 scales NU_T by
    2 and the SGS force by 4 (the model is |S| S_ij); both calls leave every
    other field alone."""
    cfg = case.channel_case(192, 129, 192, kick=0.3, streak=0.1)
    say("les_chan homogeneity", cfg.nx, cfg.ny, cfg.nz)
    g = gpu_mod.GpuChannel(cfg)
    g.initialize(*case.initial_physical(cfg))
    g.step(2)
    g.synchronize()
    st = g.state()
    h = gpu_mod.GpuChannel(cfg)
    st2 = dict(st)
    for k in ("CU1", "CU2", "CU3"):
        st2[k] = 2.0 * st[k]
    h.load_state(st2)
    n1, n2 = g.les_chan(), h.les_chan()
    e = np.abs(n2 - 2.0 * n1).max() / np.abs(n1).max()
    say("NU_T homogeneity", e)
    assert e < 1e-12, e
    a, b = g.state(), h.state()
    for k in ("CF1", "CF2", "CF3"):
        d1 = case.kept(cfg, a[k]) - case.kept(cfg, st[k])
        d2 = case.kept(cfg, b[k]) - case.kept(cfg, st2[k])
        e = np.abs(d2 - 4.0 * d1).max() / np.abs(d1).max()
        say(k, e)
        assert e < 1e-11, (k, e)
    for k in ("CP", "CTH0"):
        if k in st:
            assert np.array_equal(a[k], st[k]), k


# ---------------------------------------------------------------------------
# RK_CHAN_1 with LES: physical-space Crank-Nicolson solves with NU_T
# ---------------------------------------------------------------------------

def _les_pair(oracle_mod, gpu_mod, cfg, model=True, warm=2):
    o = oracle_mod.Oracle(cfg)
    o.initialize(*case.initial_physical(cfg))
    for _ in range(warm):
        o.step()
    g = gpu_mod.GpuChannel(cfg)
    g.load_state(o.state())
    g.time = o.time               # the oscillating pressure gradient reads TIME
    o.set_les(True, model)
    g.set_les(True, model)
    return o, g


@pytest.mark.parametrize("maker", list(CASES))
def test_les_substep_matches_oracle(oracle_mod, gpu_mod, maker):
    """Two steps of RK_CHAN_1 with the model called in every substep.  CP
    accumulates phi / H_BAR, which amplifies round-off: its bar is ten times
    the oracle's own spread when its input moves by 1e-16."""
    cfg = CASES[maker]()
    say("les substep", maker, cfg.nx, cfg.ny, cfg.nz)
    o, g = _les_pair(oracle_mod, gpu_mod, cfg)
    twin = oracle_mod.Oracle(cfg)
    rng = np.random.default_rng(5)
    twin.load_state({k: v * (1 + 1e-16 * rng.standard_normal(v.shape)) for k, v in o.state().items()})
    twin.time = o.time
    twin.set_les(True, True)
    for n in range(2):
        o.step()
        twin.step()
        g.step(1)
        ref = o.state()
        cp_sens = rel_err(cfg, twin.state()["CP"], ref["CP"])
        errs, bad = compare_states(cfg, g.state(), ref, 1e-11, fields=[f for f in FIELDS if f != "CP"])
        errs["CP"] = rel_err(cfg, g.state()["CP"], ref["CP"])
        say(n, errs, "CP self-sensitivity", cp_sens)
        assert not bad, (n, bad)
        assert errs["CP"] <= max(1e-11, 10 * cp_sens), (n, errs["CP"], cp_sens)


def test_les_spinup_is_the_dns_step(oracle_mod, gpu_mod):
    """LES on with the model not yet called (NU_T = 0): the physical-space
    solves give the DNS step of the oracle."""
    cfg = CASES["sink"]()
    say("les spin-up")
    o = oracle_mod.Oracle(cfg)
    o.initialize(*case.initial_physical(cfg))
    o.step()
    g = gpu_mod.GpuChannel(cfg)
    g.load_state(o.state())
    g.time = o.time
    g.set_les(True, model=False)
    for n in range(2):
        o.step()
        g.step(1)
    errs, bad = compare_states(cfg, g.state(), o.state(), 1e-11)
    say(errs)
    assert not bad, bad


def test_les_courant(oracle_mod, gpu_mod):
    """VARIABLE_DT with LES: COURANT on the physical velocity after the solves."""
    cfg = CASES["channel"]()
    say("les courant")
    o, g = _les_pair(oracle_mod, gpu_mod, cfg)
    for n in range(2):
        o.step(courant_cfl=0.5)
        g.step(1, courant_cfl=0.5)
        e = abs(g.delta_t - o.delta_t) / o.delta_t
        say(n, o.delta_t, e)
        assert e < 1e-12, (n, e)
    errs, bad = compare_states(cfg, g.state(), o.state(), 1e-10)
    assert not bad, bad


def test_les_refused_in_mixed_context(gpu_mod):
    cfg = CASES["channel"]()
    gm = gpu_mod.GpuChannel(cfg, precision="mixed")
    with pytest.raises(RuntimeError):
        gm.set_les(True, True)


def test_les_path_at_full_size(gpu_mod):
    """At 384x257x384: the LES path with NU_T = 0 (physical-space solves)
    against the DNS path on the same state, then two steps with the model:
    finite, and the projection leaves the velocity divergence-free."""
    cfg = case.channel_case(384, 257, 384, kick=0.3, streak=0.1)
    say("les full size", cfg.nx, cfg.ny, cfg.nz)
    a = gpu_mod.GpuChannel(cfg)
    a.initialize(*case.initial_physical(cfg))
    a.step(1)
    st = a.state()
    b = gpu_mod.GpuChannel(cfg)
    b.load_state(st)
    b.set_les(True, model=False)
    a.step(1)
    b.step(1)
    errs, bad = compare_states(cfg, b.state(), a.state(), 1e-11, fields=("CU1", "CU2", "CU3", "CP", "CTH0"))
    say(errs)
    assert not bad, bad
    b.set_les(True, model=True)
    b.step(2)
    s = b.state()
    for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
        assert np.all(np.isfinite(s[k])), k
    del a, b
    kx, kz, _, _ = case.wavenumbers(cfg)
    _, _, _, dyf = case.grid_metrics(case.grid_gy(cfg), cfg.ny)
    cu1, cu2, cu3 = (case.kept(cfg, s[k]) for k in ("CU1", "CU2", "CU3"))
    kxk = kx[:cfg.nkx + 1][None, None, :]
    kzk = kz[None, :, None]
    j = np.arange(2, cfg.ny)
    div = 1j * kxk * cu1[j] + (cu2[j + 1] - cu2[j]) / dyf[j][:, None, None] + 1j * kzk * cu3[j]
    scale = np.abs(cu2[j + 1] - cu2[j]).max() / dyf[1:cfg.ny].min() + np.abs(kxk * cu1[j]).max()
    say("div", np.abs(div).max() / scale)
    assert np.abs(div).max() < 1e-10 * scale


@pytest.mark.parametrize("maker", ["channel", "sink"])
def test_tke_les_matches_oracle(oracle_mod, gpu_mod, maker):
    """tkebudget_chan_les's epsilon_sgs after two LES steps, on planes 2..NY
    (plane 1 is NaN in both, as in the reference)."""
    cfg = CASES[maker]()
    say("tke_les", maker)
    o, g = _les_pair(oracle_mod, gpu_mod, cfg)
    for _ in range(2):
        o.step()
        g.step(1)
    eo, eg = o.tke_les(), g.tke_les()
    assert np.isnan(eo[0]) and np.isnan(eg[0])
    err = np.abs(eg[1:] - eo[1:]).max() / np.abs(eo[1:]).max()
    say("epsilon_sgs", err, np.abs(eo[1:]).max())
    assert np.abs(eo[1:]).max() > 0.0 and err < 1e-11, err

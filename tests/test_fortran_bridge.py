"""The Fortran bridge compiles against the reference's own include files
(COMMON-block layout, array shapes), so it is a drop-in for the five
subroutines.  Needs flang and the reference tree; skipped elsewhere."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from diablo_amd import case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BRIDGE = os.path.join(ROOT, "diablo_amd", "fortran", "diablo_gpu_bridge.f")
REF_FOR = "/root/reference/diablo_2.0/for"
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"


@pytest.mark.skipif(not (os.path.exists(FLANG) and os.path.isdir(REF_FOR)),
                    reason="needs flang and the reference's include files")
def test_bridge_compiles_against_reference_header():
    with tempfile.TemporaryDirectory() as d:
        # header_mpi includes mpif.h for MPI_STATUS_SIZE only; the bridge makes
        # no MPI call, so a one-parameter include is enough to check it.
        with open(os.path.join(d, "mpif.h"), "w") as f:
            f.write("      INTEGER MPI_STATUS_SIZE\n      PARAMETER (MPI_STATUS_SIZE=6)\n")
        r = subprocess.run([FLANG, "-fsyntax-only", "-ffixed-form", "-I", d, "-I", REF_FOR, BRIDGE],
                           capture_output=True, text=True, cwd=d)
        assert r.returncode == 0, r.stderr
        assert "error" not in r.stderr.lower(), r.stderr


def test_bridge_lines_fit_fixed_form():
    for n, line in enumerate(open(BRIDGE), 1):
        if line.startswith("!"):
            continue
        assert len(line.rstrip("\n")) <= 72, (n, line)


def test_bridge_defines_reference_names():
    text = open(BRIDGE).read().upper()
    for name in ("SUBROUTINE RK_CHAN_1", "SUBROUTINE REM_DIV_CHAN", "SUBROUTINE POISSON_P_CHAN",
                 "SUBROUTINE FFT_XZ_TO_FOURIER(V, VV, JMIN, JMAX)",
                 "SUBROUTINE FFT_XZ_TO_PHYSICAL(VV, V, JMIN, JMAX)"):
        assert name in text, name


# -- the bridge driven from Fortran (a program built by diablo_amd/fortran/build.py) --

def _run_file(path, cfg, nsteps, state, s1, r1):
    """The run file of dgpu_chan_run.f: parameters, grid, wavenumbers, the
    Fourier state and two transform-check fields."""
    gy_1 = case.grid_gy(cfg)
    gy, gyf, dy, dyf = case.grid_metrics(gy_1, cfg.ny)
    kx, kz, kx2, kz2 = case.wavenumbers(cfg)
    rv = [cfg.nu, cfg.lx, cfg.ly, cfg.lz, cfg.delta_t, cfg.beta, cfg.nu_v_scale, cfg.i_ro, cfg.grav_x,
          cfg.grav_y, cfg.grav_z, cfg.px0, cfg.omega0, cfg.amp_omega0, cfg.u_bc_ymin_c1, cfg.v_bc_ymin_c1,
          cfg.w_bc_ymin_c1, cfg.u_bc_ymax_c1, cfg.v_bc_ymax_c1, cfg.w_bc_ymax_c1, cfg.th_bc_ymin_c1[0],
          cfg.th_bc_ymax_c1[0], cfg.ri[0], cfg.pr[0], cfg.drhodx[0], cfg.drhodz[0]]
    iv = [cfg.f_type, cfg.u_bc_ymin, cfg.v_bc_ymin, cfg.w_bc_ymin, cfg.u_bc_ymax, cfg.v_bc_ymax,
          cfg.w_bc_ymax, cfg.th_bc_ymin[0], cfg.th_bc_ymax[0]]
    with open(path, "wb") as f:
        f.write(np.int32(nsteps).tobytes())
        f.write(np.asarray(rv, np.float64).tobytes())
        f.write(np.asarray(iv, np.int32).tobytes())
        for a in (gy, gyf, dy, dyf, kx[:cfg.nxp + 1], kx2[:cfg.nxp + 1], kz[:cfg.tnkz + 1], kz2[:cfg.tnkz + 1]):
            f.write(np.ascontiguousarray(a, np.float64).tobytes())
        for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
            f.write(np.ascontiguousarray(state[k], np.float64).tobytes())
        f.write(np.ascontiguousarray(s1).tobytes())
        f.write(np.ascontiguousarray(r1).tobytes())


def _read_out(path, cfg):
    n = int(np.prod(cfg.field_shape()))
    a = np.fromfile(path, dtype=np.float64)
    assert a.size == 11 * n + 1, a.size
    f = [a[q * n:(q + 1) * n].reshape(cfg.field_shape()) for q in range(11)]
    names = ("S1", "R1", "CU1", "CU2", "CU3", "CP", "CTH0", "CF1", "CF2", "CF3", "CFTH0")
    return dict(zip(names, f)), float(a[-1])


def _program(dims):
    from diablo_amd.fortran import build as fbuild
    return fbuild.program_path(*dims)


def _run(exe, run_path, out_path, resident):
    env = dict(os.environ)
    env["DIABLO_GPU_RESIDENT"] = "1" if resident else "0"
    print(f"running {os.path.basename(exe)} resident={resident}", flush=True)
    r = subprocess.run([exe, run_path, out_path], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and os.path.exists(out_path), r.stdout[-2000:] + r.stderr[-2000:]


@pytest.mark.gpu
def test_fortran_program_matches_oracle(oracle_mod, tmp_path):
    """The reference's call sequence from Fortran through the bridge:
    FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL of host arrays, REM_DIV_CHAN, then
    two time steps of RK_CHAN_1 (the loop of for/diablo.f:66-84) on the
    COMMON arrays, against the oracle from the same state: transforms at
    1e-13 / 1e-12, the state (CU, CP, CTH, CF, CFTH) at 1e-10; the run with
    the state kept on the GPU (DIABLO_GPU_RESIDENT=1) gives the same bytes."""
    from gpu_util import compare_states
    exe = _program((32, 33, 24))
    if not os.path.exists(exe):
        pytest.skip("the Fortran program is not built (needs flang and the reference's include files)")
    cfg = case.channel_case(32, 33, 24, kick=0.05, streak=0.05)
    o = oracle_mod.Oracle(cfg)
    o.initialize(*case.initial_physical(cfg))
    state = o.state()
    rng = np.random.default_rng(8)
    s1 = rng.standard_normal(cfg.field_shape())
    r1 = np.zeros(cfg.field_shape())
    vv = case.fourier_view(r1)
    shp = vv[:, :cfg.tnkz + 1, :cfg.nxp].shape
    vv[:, :cfg.tnkz + 1, :cfg.nxp] = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    run = str(tmp_path / "run.bin")
    _run_file(run, cfg, 2, state, s1, r1)
    outs = []
    for resident in (False, True):
        out = str(tmp_path / f"out{int(resident)}.bin")
        _run(exe, run, out, resident)
        outs.append(_read_out(out, cfg)[0])
    a, b = outs
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    s1o, r1o = s1.copy(), r1.copy()
    o.fft_array_to_fourier(s1o, 0, cfg.ny + 1)
    o.fft_array_to_physical(r1o, 0, cfg.ny + 1)
    ka, kb = case.kept(cfg, a["S1"]), case.kept(cfg, s1o)
    assert np.abs(ka - kb).max() < 1e-13 * np.abs(kb).max()
    pa, pb = a["R1"][:, :cfg.nz, :cfg.nx], r1o[:, :cfg.nz, :cfg.nx]
    assert np.abs(pa - pb).max() < 1e-12 * np.abs(pb).max()
    o.rem_div_chan()
    for _ in range(2):
        o.step()
    errs, bad = compare_states(cfg, a, o.state(), 1e-10)
    print(errs, flush=True)
    assert not bad, bad


@pytest.mark.gpu
def test_fortran_program_copy_per_step_vs_resident(tmp_path):
    """Wall time of 10 steps of the 192 x 129 x 192 channel driven from
    Fortran: the bridge's default (state copied to the host COMMON arrays
    after every step and back before the next) against
    DIABLO_GPU_RESIDENT=1 (one copy after the loop); both end on the same
    bytes.  The times are printed for the record."""
    exe = _program((192, 129, 192))
    if not os.path.exists(exe):
        pytest.skip("the Fortran program is not built (needs flang and the reference's include files)")
    from diablo_amd import gpu
    cfg = case.channel_case(192, 129, 192, re=395.0, dt=5e-4, kick=0.05, streak=0.1)
    g = gpu.GpuChannel(cfg)
    g.initialize(*case.initial_physical(cfg))
    state = g.state()
    g.close()
    z = np.zeros(cfg.field_shape())
    run = str(tmp_path / "run.bin")
    _run_file(run, cfg, 10, state, z, z)
    res = {}
    for resident in (False, True):
        out = str(tmp_path / f"out{int(resident)}.bin")
        _run(exe, run, out, resident)
        res[resident] = _read_out(out, cfg)
    (a, ta), (b, tb) = res[False], res[True]
    for k in ("CU1", "CU2", "CU3", "CP", "CTH0"):
        assert np.array_equal(a[k], b[k]), k
    pts = cfg.points * 3 * 10
    print(f"fortran 192x129x192, 10 steps: copy per step {ta / 10 * 1e3:.2f} ms/step "
          f"({pts / ta / 1e6:.0f} Mpts*substeps/s), resident {tb / 10 * 1e3:.2f} ms/step "
          f"({pts / tb / 1e6:.0f} Mpts*substeps/s)", flush=True)
    assert tb < ta

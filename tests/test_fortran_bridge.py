"""The Fortran bridge compiles against the reference's own include files
(COMMON-block layout, array shapes), so it is a drop-in for the five
subroutines.  Needs flang and the reference tree; skipped elsewhere."""
import os
import shutil
import subprocess
import tempfile

import pytest

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

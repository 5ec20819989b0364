"""Builds the Fortran test program dgpu_chan_run.f together with the bridge
(diablo_gpu_bridge.f) against the reference's own include files, and links
it with libdiablo_gpu.so:

    python -m diablo_amd.fortran.build

The reference's 'header' fixes the grid at compile time (grid_def) and the
MPI decomposition (grid_mpi); each build gets its own run configuration in
a temporary include directory: grid_def with the program's sizes, grid_mpi
for one rank, and an mpif.h carrying the single parameter the header's MPI
declarations need (neither the bridge nor the program makes an MPI call).
'header' and 'header_mpi' are read from the reference tree into that
directory for the build, because a Fortran INCLUDE looks in the including
file's own directory first: left in place they would pick the reference
tree's own grid_def and grid_mpi.
The programs go next to this file as dgpu_chan_run_<NX>x<NY>x<NZ>; they run
anywhere libdiablo_gpu.so is one directory up (rpath $ORIGIN/..).  The
build needs flang and the reference tree; without them it builds nothing.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF_FOR = "/root/reference/diablo_2.0/for"
FLANG = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
LIBDIR = os.path.dirname(HERE)

# (NX, NY, NZ): a small grid for parity with the oracle, a larger one for
# timing the bridge's host copies
SIZES = ((32, 33, 24), (192, 129, 192))


def program_path(nx: int, ny: int, nz: int) -> str:
    return os.path.join(HERE, f"dgpu_chan_run_{nx}x{ny}x{nz}")


def available() -> bool:
    return os.path.exists(FLANG) and os.path.isfile(os.path.join(REF_FOR, "header"))


def build(verbose: bool = False) -> list:
    if not available():
        if verbose:
            print("fortran: flang or the reference's include files are missing; programs not built")
        return []
    if not os.path.exists(os.path.join(LIBDIR, "libdiablo_gpu.so")):
        raise RuntimeError("build libdiablo_gpu.so first (python -m diablo_amd.build)")
    out = []
    for nx, ny, nz in SIZES:
        with tempfile.TemporaryDirectory(prefix="dgpu_fortran_") as d:
            with open(os.path.join(d, "grid_def"), "w") as f:
                f.write(f"      PARAMETER (NX={nx})\n      PARAMETER (NY={ny})\n"
                        f"      PARAMETER (NZ={nz})\n      PARAMETER (N_TH=1)\n")
            with open(os.path.join(d, "grid_mpi"), "w") as f:
                f.write("      PARAMETER(NPROCS=1)\n      PARAMETER(NPROCY=1)\n"
                        "      PARAMETER(NPROCZ=NPROCS/NPROCY)\n")
            with open(os.path.join(d, "mpif.h"), "w") as f:
                f.write("      INTEGER MPI_STATUS_SIZE\n      PARAMETER (MPI_STATUS_SIZE=6)\n")
            for name in ("header", "header_mpi"):
                shutil.copyfile(os.path.join(REF_FOR, name), os.path.join(d, name))
            exe = program_path(nx, ny, nz)
            tmp = exe + ".tmp"
            cmd = [FLANG, "-O2", "-ffixed-form", "-I", d, "-J", d,
                   os.path.join(HERE, "diablo_gpu_bridge.f"), os.path.join(HERE, "dgpu_chan_run.f"),
                   "-L", LIBDIR, "-ldiablo_gpu", "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True, cwd=d)
            if r.returncode != 0:
                raise RuntimeError("flang failed:\n" + r.stdout[-4000:] + r.stderr[-8000:])
            os.replace(tmp, exe)
            out.append(exe)
    return out


if __name__ == "__main__":
    for p in build(verbose=True):
        print(p)

"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/diablo_gpu.h declares, and the ctypes mirror of dgpu_params matches
the C struct layout (no GPU call is made)."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diablo_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dgpu_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from diablo_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        from diablo_amd import build
        build.build()
    return gpu.load_library()


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n


def test_python_mirror_lists_all_exports():
    from diablo_amd import gpu
    assert sorted(gpu.EXPORTS) == declared_functions()


def test_abi_version(lib):
    from diablo_amd import gpu
    assert lib.dgpu_abi_version() == gpu.ABI_VERSION == 2
    assert b"gfx950" in lib.dgpu_build_info()


def test_params_layout_matches_c():
    from diablo_amd import gpu
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "diablo_gpu.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dgpu_params), offsetof(dgpu_params, nu),
         offsetof(dgpu_params, th_bc_ymin_c1), offsetof(dgpu_params, drhodz),
         offsetof(dgpu_params, gy), offsetof(dgpu_params, kz2));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    P = gpu.Params
    mine = [ctypes.sizeof(P), P.nu.offset, P.th_bc_ymin_c1.offset, P.drhodz.offset, P.gy.offset, P.kz2.offset]
    assert [int(x) for x in out] == mine


def test_create_rejects_bad_arguments_without_gpu(lib):
    """Argument validation happens before any HIP call."""
    from diablo_amd import gpu
    p = gpu.Params()
    p.nx, p.ny, p.nz, p.n_th = 7, 9, 16, 1   # odd NX
    h = ctypes.c_void_p()
    assert lib.dgpu_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert not h.value

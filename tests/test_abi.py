"""The C-ABI boundary: symbols, struct layout, loud failure without the library."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from diablo3_test_amd import binding
from diablo3_test_amd.build import LIB, build_library
from helpers import ORACLE_LIB, oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diablo_gpu.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dgpu_[a-z_0-9]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def libpath():
    if not os.path.exists(LIB):
        build_library()
    return LIB


def test_library_exports_every_declared_function(libpath):
    names = declared_functions()
    assert "dgpu_rk_substep" in names and "dgpu_create" in names
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dgpu_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(libpath)  # loads without a GPU; no compute call is made
    for n in names:
        assert hasattr(lib, n)


def test_oracle_exports_same_api():
    oracle_lib()
    out = subprocess.run(["nm", "-D", "--defined-only", ORACLE_LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (orc_\w+)", out))
    # GPU-only: the timing hooks, the adjoint with its checkpoint and
    # mean-profile helpers (the adjoint's checker is the torch
    # reimplementation), the state check and the device-memory exchange
    gpu_only = ("dgpu_timing", "dgpu_kernel_class_name", "dgpu_adj_", "dgpu_mean_profile", "dgpu_checkpoint",
                "dgpu_check_state", "dgpu_set_field_device", "dgpu_get_field_device")  # also matches dgpu_check_state_noalloc
    for n in declared_functions():
        if n.startswith(gpu_only):
            continue
        assert "orc_" + n[len("dgpu_"):] in exported, n


def test_params_layout_matches_c():
    fields = [f for f, _ in binding.Params._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){",
           'printf("size %zu\\n", sizeof(dgpu_params));']
    for f in fields:
        src.append(f'printf("{f} %zu\\n", offsetof(dgpu_params, {f}));')
    src.append("return 0;}")
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write("\n".join(src))
        subprocess.run(["gcc", "-o", exe, c], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    got = dict(line.split() for line in out.strip().splitlines())
    assert int(got["size"]) == ctypes.sizeof(binding.Params)
    for f in fields:
        assert int(got[f]) == getattr(binding.Params, f).offset, f


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError, match="not found"):
        binding.load_library(str(tmp_path / "nope.so"))


def test_oracle_rejects_bad_parameters():
    from diablo3_test_amd import workloads
    from helpers import oracle_solver

    p = workloads.channel_params(48, 17, 16, 100.0)  # Nx not a power of two
    with pytest.raises(binding.DgpuError):
        oracle_solver(p, workloads.channel_grid(17))

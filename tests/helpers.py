"""Shared test helpers: the oracle's ctypes binding, solver pairs, norms."""
import ctypes
import os
import subprocess

import numpy as np

from diablo3_test_amd.binding import ChannelSolver, Params, _declare
from diablo3_test_amd import workloads

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# DGPU_ORACLE_LIB points the tests at another build of the oracle (the
# planted-mistake builds of tests/oracle_mutations.py)
ORACLE_LIB = os.environ.get("DGPU_ORACLE_LIB") or os.path.join(ORACLE_DIR, "liboracle.so")

_oracle = None


def oracle_lib():
    """The CPU oracle (oracle/liboracle.so), built here if it is missing."""
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        lib = _declare(ctypes.CDLL(ORACLE_LIB), "orc_", with_precision=False)
        P = ctypes.POINTER
        d = P(ctypes.c_double)
        for name, args in [
            ("fft_xz_to_fourier", [ctypes.c_void_p, d]),
            ("fft_xz_to_physical", [ctypes.c_void_p, d]),
            ("metrics", [ctypes.c_void_p, d, d, d, d, d]),
            ("thomas_real", [d, d, d, d, ctypes.c_int, ctypes.c_int]),
            ("thomas_complex", [d, d, d, d, ctypes.c_int, ctypes.c_int]),
        ]:
            fn = getattr(lib, "orc_" + name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _oracle = lib
    return _oracle


def oracle_solver(params, gy):
    return ChannelSolver(params, gy, lib=oracle_lib(), prefix="orc_")


def gpu_solver(params, gy, precision="fp64"):
    return ChannelSolver(params, gy, precision=precision)


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def active(c, p):
    """The dealiased block (i <= Nx/3, k <= 2 Nz/3) of a spectral array."""
    nx, nz = p["nx"], p["nz"]
    return c[..., : 2 * (nz // 3) + 1, : nx // 3 + 1]


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def field_errors(sa, sb, p, names=("cu1", "cu2", "cu3", "cth", "cp")):
    out = {}
    for k in names:
        if k in sa and k in sb:
            out[k] = rel_l2(active(sa[k], p), active(sb[k], p))
    return out


def prepared(solver, state):
    """set the state, then apply_BC_vel_mpi_post + rem_div_chan (pre_first_step)."""
    solver.set_state(state)
    solver.pre_first_step()
    return solver


def report(tag, errs):
    print(tag, " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)


def sensitivity(p, gy, st, nsteps, names=("cu1", "cu2", "cu3", "cth", "cp")):
    """Per-field relative change of the oracle's result when its input is
    perturbed at the 1e-16 level: the round-off floor that any two correct
    FP64 implementations differ by (the pressure, a small field built from
    O(amplitude^2) terms divided by h, sits far above the velocities)."""
    rng = np.random.default_rng(99)
    st2 = {k: v * (1 + 1e-16 * rng.standard_normal(v.shape)) for k, v in st.items()}
    a = prepared(oracle_solver(p, gy), st)
    b = prepared(oracle_solver(p, gy), st2)
    a.step(nsteps)
    b.step(nsteps)
    return field_errors(a.get_state(names), b.get_state(names), p, names=names)


def check_against_floor(tag, errs, floor, bar=None):
    """GPU-vs-oracle error within 50x the oracle's own round-off floor, and
    (where given) within the stated absolute bar.  The absolute bar is only
    meaningful for well-conditioned cases: with a Neumann v wall, for
    example, the oracle's 1e-16 sensitivity already reaches 2e-11 in three
    steps."""
    report(tag, errs)
    report(tag + " floor", floor)
    for k, e in errs.items():
        if bar is not None:
            assert e <= bar, (k, e)
        assert e <= 50 * floor[k] + 1e-15, (k, e, floor[k])

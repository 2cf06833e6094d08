"""The benchmark's full-size LES and mixed-precision workloads on the product
path, through the C-ABI.

* 512x257x512 AMD-Verstappen LES (the LES workload of the metric's grid):
  one substep against the oracle, and 50 steps checked through properties
  that hold at any size: a finite state, a non-negative eddy viscosity (the
  AMD closure clips its numerator at zero, les.f90:626-628, 805-809) and a
  discretely divergence-free velocity (rem_div_chan).
* 1024x385x1024 mixed precision against FP64 over 50 steps: the wall shear
  stress within 1e-5, the fields within 1e-5, the divergence at round-off.
"""
import numpy as np
import pytest

from diablo3_test_amd import stats, workloads
from helpers import field_errors, gpu_solver, oracle_solver, prepared, report
from test_gpu_mixed import compare, divergence_ratio

pytestmark = pytest.mark.gpu

SPECTRAL = ("cu1", "cu2", "cu3", "cth", "cp", "cf1", "cf2", "cf3", "cfth")


@pytest.mark.timeout(900)
def test_les_512_one_substep_vs_oracle():
    p, gy, ic, seed, amp = workloads.workload("cfg3_les_512x257x512")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    print("LES type 5 512x257x512, 1 substep vs oracle", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    g.rk_chan_1(1, 0.0, 1)
    got = g.get_state(SPECTRAL)
    nt_g = g.get_field("nu_t")
    g.close()
    o = prepared(oracle_solver(p, gy), st)
    o.rk_chan_1(1, 0.0, 1)
    errs = field_errors(got, o.get_state(SPECTRAL), p, names=SPECTRAL)
    sl = (Ellipsis, slice(0, p["nz"]), slice(0, p["nx"]))
    nt_o = o.get_field("nu_t")
    errs["nu_t"] = float(np.linalg.norm((nt_g - nt_o)[sl]) / np.linalg.norm(nt_o[sl]))
    report("les5 512x257x512 1 substep", errs)
    assert max(errs.values()) < 1e-10


@pytest.mark.timeout(900)
def test_les_512_fifty_steps_properties():
    p, gy, ic, seed, amp = workloads.workload("cfg3_les_512x257x512")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    print("LES type 5 512x257x512, 50 steps", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    for block in range(5):
        g.step(10)
        assert g.check_state(), getattr(g, "state_error", "")
        print(f"  {10 * (block + 1)} steps: finite", flush=True)
    s = g.get_state(("cu1", "cu2", "cu3"))
    div = divergence_ratio(s, p, gy)
    nt = g.get_field("nu_t")[:, : p["nz"], : p["nx"]]
    nu = 1.0 / p["re"]
    lo, hi = stats.wall_shear(s["cu1"], gy, nu)
    print(f"  divergence/term {div:.2e}  nu_t min {nt.min():.3e} max {nt.max():.3e}  "
          f"wall shear {lo:.6e} {hi:.6e}", flush=True)
    assert div < 1e-12
    assert nt.min() >= 0.0 and nt.max() > 0.0
    assert np.isfinite(lo) and np.isfinite(hi) and lo > 0 and hi > 0


@pytest.mark.timeout(1200)
def test_mixed_1024_fifty_steps_vs_fp64():
    p, gy, ic, seed, amp = workloads.workload("cfg4b_1024x385x1024")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    names = ("cu1", "cu2", "cu3", "cth", "cp")
    print("mixed vs fp64 1024x385x1024, 50 steps", flush=True)
    m = prepared(gpu_solver(p, gy, precision="mixed"), st)
    m.step(50)
    sm = m.get_state(names)
    m.close()
    print("  mixed done", flush=True)
    r = prepared(gpu_solver(p, gy), st)
    r.step(50)
    print("  fp64 done", flush=True)

    class Frozen:
        def get_state(self, names):
            return {k: sm[k] for k in names}

    e = compare("mixed 1024 50 steps", Frozen(), r, p, gy)
    assert e["wall_shear"] < 1e-5
    assert max(e[k] for k in ("cu1", "cu2", "cu3", "cth")) < 1e-5
    assert e["div_mixed"] < 1e-12

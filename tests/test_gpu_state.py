"""dgpu_check_state: the on-device finiteness scan of the state (DGPU_ESTATE)."""
import numpy as np
import pytest

from diablo3_test_amd import workloads
from helpers import gpu_solver, prepared

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_check_state_finds_injected_values(precision):
    p, gy, ic, seed, amp = workloads.workload("small_32x33x32")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    print(f"check_state {precision}", flush=True)
    g = prepared(gpu_solver(p, gy, precision=precision), st)
    assert g.check_state()
    g.step(2)
    assert g.check_state()
    good = g.get_state(("cu1", "cu2", "cu3", "cp", "cth", "cf1", "cf2", "cf3", "cfth"))
    bad = good["cu2"].copy()
    bad[5, 3, 2] = np.nan
    g.set_field("cu2", bad)
    assert not g.check_state()
    assert "cu2" in g.state_error and "cu1" not in g.state_error
    g.set_field("cu2", good["cu2"])
    assert g.check_state()
    badf = good["cfth"].copy()
    badf[0, 7, 1, 4] = np.inf
    g.set_field("cfth", badf)
    assert not g.check_state() and "cfth" in g.state_error
    g.set_field("cfth", good["cfth"])
    # a non-finite value spreads through a step to every velocity field
    g.set_field("cu2", bad)
    g.step(1)
    assert not g.check_state()
    for k in ("cu1", "cu2", "cu3", "cp"):
        assert k in g.state_error


def test_check_state_les_eddy_viscosity():
    p = workloads.channel_params(32, 33, 32, 500.0, ri=0.1, dt=1e-3, les=True)
    gy = workloads.channel_grid(33)
    st = workloads.initial_state(p, gy, "vonkarman", 3, 0.1)
    for precision in ("fp64", "mixed"):
        print(f"check_state LES {precision}", flush=True)
        g = prepared(gpu_solver(p, gy, precision=precision), st)
        g.step(1)
        assert g.check_state()
        nt = g.get_field("nu_t")
        nt[10, 4, 6] = np.nan
        g.set_field("nu_t", nt)
        assert not g.check_state() and "nu_t" in g.state_error

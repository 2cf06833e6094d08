"""dgpu_check_state_noalloc against dgpu_check_state through the C-ABI.

Two calls in a row on two handles, the second on a smaller grid, each
with a finite state and with an injected NaN: the no-allocation entry
point returns the same code and names the same fields as the entry point
that allocates its flags.
"""
import numpy as np
import pytest

from diablo3_test_amd import workloads
from helpers import gpu_solver, prepared

pytestmark = pytest.mark.gpu


def both(g):
    a = g.check_state()
    ea = getattr(g, "state_error", None) if not a else None
    b = g.check_state_noalloc()
    eb = getattr(g, "state_error", None) if not b else None
    return (a, ea), (b, eb)


def test_noalloc_check_matches_allocating_check():
    cases = [(64, 65, 64), (32, 33, 32)]  # the second call's input is smaller
    for nx, ny, nz in cases:
        p = workloads.channel_params(nx, ny, nz, 180.0, ri=0.1, dt=1e-3)
        gy = workloads.channel_grid(ny)
        st = workloads.initial_state(p, gy, "white", 4, 1e-2)
        print(f"noalloc check {nx}x{ny}x{nz}", flush=True)
        g = prepared(gpu_solver(p, gy), st)
        g.step(1)
        r_alloc, r_noalloc = both(g)
        assert r_alloc == r_noalloc == (True, None)
        # the no-allocation call first, then the allocating one, on a bad state
        bad = g.get_field("cu3")
        bad[ny // 2, 2, 1] = np.nan
        g.set_field("cu3", bad)
        b = g.check_state_noalloc()
        eb = g.state_error
        a = g.check_state()
        ea = g.state_error
        assert (a, ea) == (b, eb) and not a and "cu3" in ea
        # and a clean state again: the scratch keeps the last call's flags,
        # which the next scan overwrites
        bad[ny // 2, 2, 1] = 0.0
        g.set_field("cu3", bad)
        assert g.check_state_noalloc() and g.check_state()
        g.close()

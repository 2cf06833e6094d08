"""The HIP substep with the subgrid models on (use_LES, for/les.f90) against
the oracle on the same seeded states, through the C-ABI."""
import numpy as np
import pytest

from diablo3_test_amd import workloads
from helpers import (active, check_against_floor, field_errors, gpu_solver, oracle_solver, prepared,
                     rel_l2, report, sensitivity)

pytestmark = pytest.mark.gpu

SPECTRAL = ("cu1", "cu2", "cu3", "cth", "cp", "cf1", "cf2", "cf3", "cfth")


def les_case(typ, nx=32, ny=33, nz=32, **kw):
    p = workloads.channel_params(nx, ny, nz, 2000.0, ri=0.1, dt=1e-3, les=True, les_model_type=typ)
    p.update(kw)
    gy = workloads.channel_grid(ny)
    st = workloads.initial_state(p, gy, "vonkarman", 3, 1e-1)
    fluct = st["cu1"][2:ny].copy()
    fluct[:, 0, 0] = 0.0
    for n in range(p["n_th"]):
        st["cth"][n, 2:ny] += 0.5 * fluct  # scalar fluctuations for the scalar closures
    return p, gy, st


def physical_err(a, b, p):
    sl = (Ellipsis, slice(0, p["nz"]), slice(0, p["nx"]))
    return rel_l2(a[sl], b[sl])


def run_pair(p, gy, st, nsteps):
    g = prepared(gpu_solver(p, gy), st)
    o = prepared(oracle_solver(p, gy), st)
    g.step(nsteps)
    o.step(nsteps)
    return g, o


@pytest.mark.parametrize("typ", [1, 4, 5])
def test_les_three_steps(typ):
    p, gy, st = les_case(typ)
    print(f"LES type {typ} 32x33x32, 3 steps", flush=True)
    g, o = run_pair(p, gy, st, 3)
    errs = field_errors(g.get_state(SPECTRAL), o.get_state(SPECTRAL), p, names=SPECTRAL)
    floor = sensitivity(p, gy, st, 3, names=SPECTRAL)
    check_against_floor(f"les{typ}", errs, floor)
    for name in ("nu_t", "kappa_t"):
        e = physical_err(g.get_field(name), o.get_field(name), p)
        print(name, f"{e:.2e}", flush=True)
        assert e < 1e-10
    assert np.abs(g.get_field("nu_t")).max() > 1e-4


NEUMANN = dict(u_bc_ymin=1, w_bc_ymin=1, u_bc_ymin_c1=0.05, w_bc_ymin_c1=-0.02,
               u_bc_ymax=1, w_bc_ymax=1, u_bc_ymax_c1=0.0, w_bc_ymax_c1=0.01,
               th_bc_ymin=[1], th_bc_ymin_c1=[0.3], th_bc_ymax=[1], th_bc_ymax_c1=[-0.1])


@pytest.mark.parametrize("typ", [1, 5])
def test_les_neumann_walls(typ):
    """Stress-free / flux walls: jstart = 1, the ghost copies of apply_BC_les
    and apply_BC_th_les, and the scalar wall values of les_chan."""
    p, gy, st = les_case(typ, **NEUMANN)
    print(f"LES type {typ} neumann walls, 3 steps", flush=True)
    g, o = run_pair(p, gy, st, 3)
    errs = field_errors(g.get_state(SPECTRAL), o.get_state(SPECTRAL), p, names=SPECTRAL)
    floor = sensitivity(p, gy, st, 3, names=SPECTRAL)
    check_against_floor(f"les{typ} neumann", errs, floor)


def test_les_rectangular_two_scalars():
    p, gy, st = les_case(5, nx=64, ny=41, nz=16, n_th=2, pr=[1.0, 0.7], th_bc_ymin=[0, 1],
                         th_bc_ymax=[0, 0], th_bc_ymin_c1=[-1.0, 0.2], th_bc_ymax_c1=[1.0, 0.5],
                         dthdx=[0.0, 0.1], dthdz=[0.0, 0.0])
    print("LES type 5 64x41x16 two scalars, 2 steps", flush=True)
    g, o = run_pair(p, gy, st, 2)
    errs = field_errors(g.get_state(SPECTRAL), o.get_state(SPECTRAL), p, names=SPECTRAL)
    floor = sensitivity(p, gy, st, 2, names=SPECTRAL)
    check_against_floor("les5 64x41x16 nth2", errs, floor)
    e = physical_err(g.get_field("kappa_t"), o.get_field("kappa_t"), p)
    assert e < 1e-10


def test_les_startup_gate():
    """create_new_flow: model off for time steps <= 100 (channel.f90:259)."""
    p, gy, st = les_case(5, create_new_flow=1)
    print("LES gate", flush=True)
    res = {}
    for ts in (100, 101):
        g = prepared(gpu_solver(p, gy), st)
        o = prepared(oracle_solver(p, gy), st)
        g.rk_chan_1(1, 0.0, ts)
        o.rk_chan_1(1, 0.0, ts)
        errs = field_errors(g.get_state(SPECTRAL), o.get_state(SPECTRAL), p, names=SPECTRAL)
        report(f"gate ts={ts}", errs)
        assert max(errs.values()) < 1e-11
        res[ts] = g.get_state(("cf1",))["cf1"]
    assert rel_l2(active(res[101], p), active(res[100], p)) > 1e-6


def test_les_cfg3_shape_one_substep():
    """The 512x257x512 LES workload's grid at reduced horizontal size
    (128x257x128): one substep against the oracle."""
    p = workloads.channel_params(128, 257, 128, 2000.0, ri=0.1, dt=1e-3, les=True, les_model_type=5)
    gy = workloads.channel_grid(257)
    st = workloads.initial_state(p, gy, "vonkarman", 3, 1e-1)
    print("LES type 5 128x257x128, 1 substep", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    o = prepared(oracle_solver(p, gy), st)
    g.rk_chan_1(1, 0.0, 1)
    o.rk_chan_1(1, 0.0, 1)
    errs = field_errors(g.get_state(SPECTRAL), o.get_state(SPECTRAL), p, names=SPECTRAL)
    report("les5 128x257x128", errs)
    assert max(errs.values()) < 1e-10


@pytest.mark.parametrize("ny", [289, 385])
def test_les_tall_channel(ny):
    """The LES CN solve (and the whole LES substep) on columns of 291 and 387
    rows against the oracle, FP64 and the mixed form."""
    p, gy, st = les_case(5, nx=16, ny=ny, nz=16)
    print(f"LES type 5 16x{ny}x16, 2 steps", flush=True)
    o = prepared(oracle_solver(p, gy), st)
    o.step(2)
    want = o.get_state(SPECTRAL)
    for prec in ("fp64", "mixed"):
        g = prepared(gpu_solver(p, gy, precision=prec), st)
        g.step(2)
        errs = field_errors(g.get_state(SPECTRAL), want, p, names=SPECTRAL)
        g.close()
        report(f"les {prec} Ny={ny}", errs)
        state = 1e-12 if prec == "fp64" else 1e-6
        for k, e in errs.items():
            bar = (1e-9 if prec == "fp64" else 1e-4) if k == "cp" else \
                  (1e-10 if prec == "fp64" else 1e-4) if k.startswith("cf") else state
            assert e < bar, (prec, k, e)

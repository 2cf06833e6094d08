"""Many passive scalars, and long z transforms whose RHS pass does not fit in
one launch's LDS, against the oracle through the C-ABI.

* N_th = 6 and N_th = 16 (the cap): each scalar with its own Prandtl number,
  background gradients, wall conditions and buoyancy share, so that a scalar
  read from the wrong slot or row range shows up in its own field.
* Nz = 1024 with three scalars and Nz = 2048 with two: the product spectra of
  a substep exceed the 160 KiB of LDS a workgroup has, and the RHS pass runs
  as several launches with disjoint outputs.
"""
import numpy as np
import pytest

from diablo3_test_amd import workloads
from helpers import field_errors, gpu_solver, oracle_solver, prepared, report

pytestmark = pytest.mark.gpu


def scalar_params(nx, ny, nz, n_th, re=400.0, dt=5e-4):
    p = workloads.channel_params(nx, ny, nz, re, ri=0.0, dt=dt, n_th=n_th)
    rng = np.random.default_rng(100 + n_th)
    p["pr"] = [float(x) for x in rng.uniform(0.5, 2.0, n_th)]
    p["dthdx"] = [float(x) for x in rng.uniform(-0.2, 0.2, n_th)]
    p["dthdz"] = [float(x) for x in rng.uniform(-0.2, 0.2, n_th)]
    p["th_bc_ymin"] = [n % 2 for n in range(n_th)]
    p["th_bc_ymax"] = [(n // 2) % 2 for n in range(n_th)]
    p["th_bc_ymin_c1"] = [float(x) for x in rng.uniform(-1.0, 1.0, n_th)]
    p["th_bc_ymax_c1"] = [float(x) for x in rng.uniform(-1.0, 1.0, n_th)]
    p["grav_y"] = 0.05
    p["grav_x"] = 0.01
    return p


def scalar_state(p, gy, seed):
    """White-noise velocities and one independent noise field per scalar."""
    st = workloads.initial_state(p, gy, "white", seed, 1e-2)
    rng = np.random.default_rng(seed + 1)
    cth = st["cth"]
    for n in range(p["n_th"]):
        cth[n] = cth[n] * rng.uniform(0.5, 1.5) + 0.01 * n * cth[0]
    st["cth"] = cth
    return st


def run_pair(p, gy, st, nsteps, precision="fp64"):
    g = prepared(gpu_solver(p, gy, precision=precision), st)
    o = prepared(oracle_solver(p, gy), st)
    g.step(nsteps)
    o.step(nsteps)
    names = ("cu1", "cu2", "cu3", "cp", "cth", "cfth")
    errs = field_errors(g.get_state(names), o.get_state(names), p, names=names)
    # per scalar, so that one scalar's mistake is not averaged away
    gt, ot = g.get_field("cth"), o.get_field("cth")
    for n in range(p["n_th"]):
        errs[f"th{n}"] = float(np.linalg.norm(gt[n] - ot[n]) / np.linalg.norm(ot[n]))
    g.close()
    return errs


# relative L2 bars: state fields, the pressure, the scalars' RK carry.  The
# carry is the explicit right-hand side, dominated by the diffusion of the
# highest wavenumbers (kz^2 ~ 5e5 at Nz = 2048), so it carries the transforms'
# round-off amplified; in the mixed form it is stored in FP32.
BARS = {"fp64": (1e-12, 1e-9, 1e-10), "mixed": (1e-6, 1e-4, 3e-6)}


def check(errs, precision):
    state, cp, carry = BARS[precision]
    assert max(v for k, v in errs.items() if k not in ("cp", "cfth")) < state, errs
    assert errs["cp"] < cp, errs
    assert errs["cfth"] < carry, errs


@pytest.mark.parametrize("n_th,shape,nsteps", [(6, (32, 33, 32), 3), (16, (16, 17, 16), 2)])
def test_many_scalars_vs_oracle(n_th, shape, nsteps):
    nx, ny, nz = shape
    p = scalar_params(nx, ny, nz, n_th)
    gy = workloads.channel_grid(ny)
    st = scalar_state(p, gy, 21)
    print(f"N_th={n_th} {nx}x{ny}x{nz}, {nsteps} steps vs oracle", flush=True)
    errs = run_pair(p, gy, st, nsteps)
    report(f"nth{n_th}", errs)
    check(errs, "fp64")


@pytest.mark.parametrize("model", [4, 5])
def test_many_scalars_les_vs_oracle(model):
    """LES with six scalars: one kappa_t per scalar (les_chan_th, AMD and
    AMD-Verstappen scalar models), the gate open (a restart)."""
    p = scalar_params(32, 33, 32, 6)
    p["use_les"] = 1
    p["les_model_type"] = model
    gy = workloads.channel_grid(33)
    st = scalar_state(p, gy, 24)
    print(f"LES type {model} N_th=6 32x33x32, 2 steps vs oracle", flush=True)
    errs = run_pair(p, gy, st, 2)
    report(f"les{model} nth6", errs)
    check(errs, "fp64")


def test_many_scalars_mixed_vs_oracle():
    p = scalar_params(32, 33, 32, 6)
    gy = workloads.channel_grid(33)
    st = scalar_state(p, gy, 22)
    print("mixed N_th=6 32x33x32, 3 steps vs oracle", flush=True)
    errs = run_pair(p, gy, st, 3, precision="mixed")
    report("mixed nth6", errs)
    check(errs, "mixed")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("n_th,shape", [(3, (16, 17, 1024)), (2, (16, 9, 2048))])
def test_split_rhs_pass_vs_oracle(n_th, shape, precision):
    nx, ny, nz = shape
    p = scalar_params(nx, ny, nz, n_th)
    gy = workloads.channel_grid(ny)
    st = scalar_state(p, gy, 23)
    print(f"{precision} N_th={n_th} {nx}x{ny}x{nz}, 1 step vs oracle", flush=True)
    errs = run_pair(p, gy, st, 1, precision=precision)
    report(f"{precision} nth{n_th} {nz}", errs)
    check(errs, precision)

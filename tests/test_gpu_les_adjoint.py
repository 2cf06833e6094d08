"""The adjoint of the FP64 LES substep (subgrid models 1, 4, 5, the CN solve
with the eddy viscosity, the wall values of the LES stage) against central
finite differences of the GPU forward substep, and the checkpointed
wall-shear gradient through LES steps against finite differences.

It also checks the adjoint against torch autograd through the FP64 torch
restatement of the LES substep (tests/torch_les.py, pinned to the oracle in
tests/test_torch_step.py).

The directions are physical: random real fields transformed to spectral
space (so the kx = 0 modes are Hermitian).  The bar is 1e-6 relative, the
bar for the adjoint gradient.  The AMD clip max(0, .) makes models 4 and 5
only piecewise smooth, so a central difference is exact only while no clip
switches inside +-eps: for one substep eps = 1e-7 on O(1) fields keeps the
differences on one smooth piece (a scan over eps = 1e-4 .. 1e-8 agrees to
<= 5e-10 at 1e-7 and to ~1e-9 at 1e-8, where roundoff takes over, while
1e-6 crossed a clip by 2e-6 in one direction).  The wall-shear directions
carry a mean (kx = kz = 0) part, so that the gradient along them is well
above the differences' roundoff."""
import numpy as np
import pytest
import torch

from diablo3_test_amd import workloads
from diablo3_test_amd.adjoint import STATE, WallShearRun
from helpers import active, gpu_solver, oracle_solver, prepared, rel_l2, report
from torch_step import TorchChannel, from_ref, to_ref

pytestmark = pytest.mark.gpu


def les_case(model, nx=32, ny=33, nz=32, **kw):
    p = workloads.channel_params(nx, ny, nz, 2000.0, ri=0.1, dt=1e-3, les=True, les_model_type=model)
    p.update(kw)
    gy = workloads.channel_grid(ny)
    st = workloads.initial_state(p, gy, "vonkarman", 3, 0.3)
    return p, gy, st


def physical_direction(p, gy, seed, fields=STATE, mean=False):
    """Random smooth physical fields on every state field, as spectral arrays;
    mean True adds a random y profile (the kx = kz = 0 mode) to each."""
    rng = np.random.default_rng(seed)
    o = oracle_solver(p, gy)
    nx, ny, nz, nth = p["nx"], p["ny"], p["nz"], p["n_th"]
    x = np.arange(nx) / nx
    z = np.arange(nz) / nz
    y = np.linspace(-1.0, 1.0, ny + 2)
    out = {}
    for k in fields:
        base = "cu1" if k in ("cf1", "cf2", "cf3", "cp") else ("cth" if k == "cfth" else k)
        lead = (nth,) if base == "cth" else ()
        f = np.zeros(lead + (ny + 2, nz + 2, nx + 2))
        for _ in range(6):
            kx, kz = rng.integers(0, 4), rng.integers(0, 4)
            ph1, ph2 = rng.uniform(0, 2 * np.pi, 2)
            amp = rng.standard_normal(lead + (ny + 2, 1, 1)) * (1 - y[:, None, None] ** 2)
            f[..., :nz, :nx] += amp * (np.cos(2 * np.pi * kx * x[None, None, :] + ph1) *
                                       np.cos(2 * np.pi * kz * z[None, :, None] + ph2))
        if mean:
            f[..., :nz, :nx] += rng.standard_normal(lead + (ny + 2, 1, 1)) * (1 - y[:, None, None] ** 2)
        o.set_physical(base, f)
        out[k] = o.get_field(base)
    return out


def loss(g, p, s, lam, rk, t, ts):
    g.set_state(s)
    g.rk_chan_1(rk, t, ts)
    out = g.get_state(STATE)
    return sum(np.sum((np.conj(active(lam[k], p)) * active(out[k], p)).real) for k in STATE)


@pytest.mark.parametrize("model,variant", [(1, "dirichlet"), (4, "dirichlet"), (5, "dirichlet"),
                                           (5, "neumann"), (1, "neumann")])
@pytest.mark.parametrize("rk", [1, 3])
def test_les_adjoint_substep_finite_differences(model, variant, rk):
    extra = {}
    if variant == "neumann":
        extra = dict(u_bc_ymin=1, w_bc_ymax=1, u_bc_ymin_c1=0.1, w_bc_ymax_c1=0.02, th_bc_ymax=[1],
                     th_bc_ymax_c1=[0.2], v_bc_ymax=1)
    p, gy, st = les_case(model, **extra)
    print(f"LES adjoint substep model {model} {variant} rk={rk}", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    g.step(2)  # a state with a live RK carry and pressure
    s0 = g.get_state(STATE)
    t, ts = 0.3, 7
    lam = physical_direction(p, gy, 900 + rk)
    g.set_state(s0)
    g.adj_set_state(lam)
    g.adj_rk_chan_1(rk, t, ts)
    adj = g.adj_get_state(STATE)
    # the forward state is left as it was
    back = g.get_state(STATE)
    for k in STATE:
        np.testing.assert_array_equal(active(back[k], p), active(s0[k], p))
    eps = 1e-7
    worst = 0.0
    for i in range(3):
        d = physical_direction(p, gy, 1000 + 10 * rk + i)
        ad = sum(np.sum((np.conj(adj[k]) * d[k]).real) for k in STATE)
        lp = loss(g, p, {k: s0[k] + eps * d[k] for k in STATE}, lam, rk, t, ts)
        lm = loss(g, p, {k: s0[k] - eps * d[k] for k in STATE}, lam, rk, t, ts)
        fd = (lp - lm) / (2 * eps)
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"  direction {i}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
    g.close()
    assert worst < 1e-6


def test_les_wall_shear_gradient_finite_differences():
    """AMD-Verstappen LES, 32x33x32, 6 steps, checkpoints every 3: <grad, d>
    against central differences along 4 divergence-free directions."""
    p, gy, st = les_case(5)
    g = prepared(gpu_solver(p, gy), st)
    s0 = g.get_state(STATE)
    print("LES wall-shear gradient: 6 steps, checkpoints every 3", flush=True)
    run = WallShearRun(g, 6, checkpoint_every=3, time0=0.0, time_step0=200)
    L, grad = run.gradient(s0)
    print("loss", L, flush=True)
    eps = 1e-6
    worst = 0.0
    for i in range(4):
        d = physical_direction(p, gy, 2000 + i, fields=("cu1", "cu2", "cu3"), mean=True)
        # project the velocity direction: divergence-free, wall values applied
        h = gpu_solver(p, gy)
        h.set_state({k: d[k] for k in ("cu1", "cu2", "cu3")})
        h.pre_first_step()
        dv = {k: h.get_field(k) for k in ("cu1", "cu2", "cu3")}
        h.close()
        ad = sum(np.sum((np.conj(grad[k]) * dv[k]).real) for k in dv)
        lp = run.loss({**s0, **{k: s0[k] + eps * dv[k] for k in dv}})
        lm = run.loss({**s0, **{k: s0[k] - eps * dv[k] for k in dv}})
        fd = (lp - lm) / (2 * eps)
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"  direction {i}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
    assert worst < 1e-6


TORCH_VARIANTS = {
    "dirichlet": {},
    "neumann": dict(u_bc_ymin=1, w_bc_ymax=1, u_bc_ymin_c1=0.1, w_bc_ymax_c1=0.02, th_bc_ymax=[1],
                    th_bc_ymax_c1=[0.2], v_bc_ymax=1),
}


@pytest.mark.parametrize("model", [1, 4, 5])
@pytest.mark.parametrize("variant", sorted(TORCH_VARIANTS))
@pytest.mark.parametrize("rk", [1, 2, 3])
def test_les_adjoint_substep_matches_torch(model, variant, rk):
    """One LES substep's vector-Jacobian product, 16x17x16, gate open: the HIP
    adjoint against torch autograd of the torch restatement, every state
    field within 1e-10 relative L2 (cp: 1e-8, the pressure's round-off
    floor, as in the DNS check)."""
    p, gy, st = les_case(model, 16, 17, 16, **TORCH_VARIANTS[variant])
    o = prepared(oracle_solver(p, gy), st)
    o.step(2)  # a live RK carry, pressure and eddy viscosity
    s0 = o.get_state(STATE)
    lam = physical_direction(p, gy, 4000 + rk)
    rng = np.random.default_rng(50 + rk)
    for k in STATE:  # plus a random non-smooth part on every active mode
        a = active(lam[k], p)
        a += 0.1 * (rng.standard_normal(a.shape) + 1j * rng.standard_normal(a.shape))
    print(f"LES adjoint vs torch: model {model} {variant} rk={rk}", flush=True)
    g = gpu_solver(p, gy)
    g.set_state(s0)
    g.adj_set_state(lam)
    g.adj_rk_chan_1(rk, 0.3, 7)
    got = g.adj_get_state(STATE)
    g.close()
    tc = TorchChannel(p, gy)
    t = from_ref(s0, p)
    for v in t.values():
        v.requires_grad_(True)
    out = tc.substep(t, rk, 0.3, 7)
    lt = from_ref(lam, p)
    L = sum((torch.conj(lt[k]) * out[k]).real.sum() for k in STATE)
    L.backward()
    want = to_ref({k: t[k].grad for k in STATE}, p)
    errs = {k: rel_l2(active(got[k], p), active(want[k], p)) for k in STATE}
    report(f"LES adj vs torch m{model} {variant} rk{rk}", errs)
    for k, e in errs.items():
        assert np.isfinite(e) and e < (1e-8 if k == "cp" else 1e-10), (k, e)

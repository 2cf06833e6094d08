"""The hand-written HIP adjoint of the FP64 DNS substep against torch autograd
through the FP64 torch reimplementation (tests/torch_step.py, itself checked
against the oracle), and the checkpointed wall-shear gradient against central
finite differences of the GPU forward run."""
import numpy as np
import pytest
import torch

from diablo3_test_amd import workloads
from diablo3_test_amd.adjoint import STATE, WallShearLoss, WallShearRun, wall_shear_weights
from helpers import active, gpu_solver, oracle_solver, prepared, rel_l2, report
from torch_step import TorchChannel, from_ref, to_ref

pytestmark = pytest.mark.gpu


def seeded(p, gy, seed=7, amp=1e-2, carry=True):
    st = workloads.initial_state(p, gy, "white", seed, amp)
    o = prepared(oracle_solver(p, gy), st)
    s = o.get_state(STATE)
    if carry:
        rng = np.random.default_rng(seed + 1)
        for k in ("cf1", "cf2", "cf3", "cfth"):
            s[k] = s[k] + 1e-3 * (rng.standard_normal(s[k].shape) + 1j * rng.standard_normal(s[k].shape))
    return s


def random_adjoint(p, seed):
    rng = np.random.default_rng(seed)
    shp = (p["ny"] + 2, p["nz"] + 2, p["nx"] // 2 + 1)
    nkx, nkz = p["nx"] // 3, p["nz"] // 3
    out = {}
    for k in STATE:
        s = ((p["n_th"],) + shp) if k in ("cth", "cfth") else shp
        a = np.zeros(s, dtype=np.complex128)
        blk = (Ellipsis, slice(0, 2 * nkz + 1), slice(0, nkx + 1))
        a[blk] = rng.standard_normal(a[blk].shape) + 1j * rng.standard_normal(a[blk].shape)
        out[k] = a
    return out


def torch_vjp(p, gy, s_ref, lam_ref, rk, time):
    tc = TorchChannel(p, gy)
    s = from_ref(s_ref, p)
    for v in s.values():
        v.requires_grad_(True)
    out = tc.substep(s, rk, time)
    lam = from_ref(lam_ref, p)
    L = sum((torch.conj(lam[k]) * out[k]).real.sum() for k in STATE)
    L.backward()
    return to_ref({k: s[k].grad for k in STATE}, p)


VARIANTS = {
    "default": {},
    "neumann": dict(u_bc_ymin=1, w_bc_ymax=1, u_bc_ymin_c1=0.1, w_bc_ymax_c1=0.02, th_bc_ymax=[1],
                    th_bc_ymax_c1=[0.2], v_bc_ymax=1),
    "coriolis_gravity": dict(ro_inv=0.5, grav_x=0.2, grav_z=-0.1, grav_y=0.4, dthdx=[0.1], dthdz=[0.05],
                             pr=[0.7]),
    "front": dict(flavor_front=1, ly=1.5, ro_inv=0.5, delta=1.5, dthdx=[0.2], dthdz=[-0.1], pr=[0.7]),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
@pytest.mark.parametrize("rk", [1, 2, 3])
def test_adjoint_substep_matches_torch(name, rk):
    p = workloads.channel_params(16, 17, 16, 200.0, ri=0.1, dt=2e-3)
    p.update(VARIANTS[name])
    gy = workloads.channel_grid(p["ny"])
    s = seeded(p, gy)
    lam = random_adjoint(p, 40 + rk)
    print(f"adjoint substep {name} rk={rk}", flush=True)
    g = gpu_solver(p, gy)
    g.set_state(s)
    g.adj_set_state(lam)
    g.adj_rk_chan_1(rk, 0.3, 5)
    got = g.adj_get_state(STATE)
    # the forward state is left as it was
    back = g.get_state(STATE)
    for k in STATE:
        np.testing.assert_array_equal(active(back[k], p), active(s[k], p))
    want = torch_vjp(p, gy, s, lam, rk, 0.3)
    errs = {k: rel_l2(active(got[k], p), active(want[k], p)) for k in STATE}
    report(f"adj {name} rk{rk}", errs)
    for k, e in errs.items():
        assert e < 1e-11, (k, e)


@pytest.mark.parametrize("ny", [289, 545])
def test_adjoint_substep_tall_channel(ny):
    """Ny past 257 and past 513: the transposed projection's 8-row segment
    layouts against torch autograd."""
    p = workloads.channel_params(16, ny, 16, 200.0, ri=0.1, dt=2e-4)
    gy = workloads.channel_grid(p["ny"])
    s = seeded(p, gy)
    lam = random_adjoint(p, 77)
    print(f"adjoint substep 16x{ny}x16 rk=2", flush=True)
    g = gpu_solver(p, gy)
    g.set_state(s)
    g.adj_set_state(lam)
    g.adj_rk_chan_1(2, 0.3, 5)
    got = g.adj_get_state(STATE)
    want = torch_vjp(p, gy, s, lam, 2, 0.3)
    errs = {k: rel_l2(active(got[k], p), active(want[k], p)) for k in STATE}
    report(f"adj 16x{ny}x16", errs)
    for k, e in errs.items():
        assert e < 1e-11, (k, e)


@pytest.mark.parametrize("rk", [1, 3])
def test_adjoint_substep_nx512_matches_torch(rk):
    """Nx = 512: the 512-point x transforms of the forward and adjoint passes
    (the one-wavefront cross-lane form) against torch autograd, 512x17x16."""
    p = workloads.channel_params(512, 17, 16, 200.0, ri=0.1, dt=2e-3)
    gy = workloads.channel_grid(p["ny"])
    s = seeded(p, gy)
    lam = random_adjoint(p, 60 + rk)
    print(f"adjoint substep 512x17x16 rk={rk}", flush=True)
    g = gpu_solver(p, gy)
    g.set_state(s)
    g.adj_set_state(lam)
    g.adj_rk_chan_1(rk, 0.3, 5)
    got = g.adj_get_state(STATE)
    want = torch_vjp(p, gy, s, lam, rk, 0.3)
    errs = {k: rel_l2(active(got[k], p), active(want[k], p)) for k in STATE}
    report(f"adj 512x17x16 rk{rk}", errs)
    for k, e in errs.items():
        assert e < 1e-11, (k, e)


def test_wall_shear_gradient_matches_torch():
    """L over 3 steps with checkpoints every 2: HIP checkpointed adjoint vs torch autograd."""
    p = workloads.channel_params(16, 17, 16, 200.0, ri=0.1, dt=2e-3)
    gy = workloads.channel_grid(p["ny"])
    s = seeded(p, gy, carry=False)
    nsteps = 3
    print("wall-shear gradient 16x17x16, 3 steps", flush=True)
    run = WallShearRun(gpu_solver(p, gy), nsteps, checkpoint_every=2)
    L, grad = run.gradient(s)
    tc = TorchChannel(p, gy)
    t = from_ref(s, p)
    for v in t.values():
        v.requires_grad_(True)
    w = torch.tensor(wall_shear_weights(p, gy))
    cur, Lt = t, 0.0
    for n in range(nsteps):
        cur = tc.step(cur, 0.0)
        Lt = Lt + (w * cur["cu1"][:, 0, 0].real).sum() / nsteps
    Lt.backward()
    want = to_ref({k: t[k].grad for k in STATE}, p)
    print("loss", L, float(Lt.detach()), flush=True)
    assert abs(L - float(Lt)) <= 1e-12 * abs(L)
    errs = {k: rel_l2(active(grad[k], p), active(want[k], p)) for k in STATE}
    report("grad vs torch", errs)
    for k, e in errs.items():
        # the gradient with respect to the initial pressure reaches the loss
        # only through -h grad(cp) of the first substep and sits at the same
        # round-off floor (~1e-9 relative) as the forward pressure
        assert e < (1e-8 if k == "cp" else 1e-10), (k, e)


def test_wall_shear_gradient_matches_torch_32():
    """The size the cfg5 check names for the torch reimplementation:
    32x33x32, Re 400, 10 steps with a checkpoint every 5 (the recomputing
    cadence of the 50-step run, scaled), HIP checkpointed adjoint against
    torch autograd through all 30 substeps."""
    p = workloads.channel_params(32, 33, 32, 400.0, ri=0.1, dt=1e-3)
    gy = workloads.channel_grid(p["ny"])
    s = seeded(p, gy, amp=5e-2, carry=False)
    nsteps = 10
    print("wall-shear gradient 32x33x32, 10 steps, checkpoints every 5", flush=True)
    run = WallShearRun(gpu_solver(p, gy), nsteps, checkpoint_every=5)
    L, grad = run.gradient(s)
    tc = TorchChannel(p, gy)
    t = from_ref(s, p)
    for v in t.values():
        v.requires_grad_(True)
    w = torch.tensor(wall_shear_weights(p, gy))
    cur, Lt = t, 0.0
    for n in range(nsteps):
        cur = tc.step(cur, 0.0)
        Lt = Lt + (w * cur["cu1"][:, 0, 0].real).sum() / nsteps
    Lt.backward()
    want = to_ref({k: t[k].grad for k in STATE}, p)
    print("loss", L, float(Lt.detach()), flush=True)
    assert abs(L - float(Lt.detach())) <= 1e-12 * abs(L)
    errs = {k: rel_l2(active(grad[k], p), active(want[k], p)) for k in STATE}
    report("grad vs torch 32x33x32", errs)
    for k, e in errs.items():
        assert e < (1e-8 if k == "cp" else 1e-10), (k, e)


def divergence_free_direction(p, gy, seed):
    """A random velocity perturbation projected with apply_BC_vel_mpi_post +
    rem_div_chan, scaled to unit RMS over the active modes; other fields zero."""
    st = workloads.initial_state(p, gy, "powerlaw", seed, 1.0)
    st["cu1"][:, 0, 0] = 0.0  # perturb the fluctuations, not the mean profile
    g = prepared(gpu_solver(p, gy), st)
    d = {k: np.zeros_like(v) for k, v in g.get_state(STATE).items()}
    for k in ("cu1", "cu2", "cu3"):
        d[k] = g.get_field(k)
    rms = np.sqrt(sum(np.sum(np.abs(active(d[k], p)) ** 2) for k in ("cu1", "cu2", "cu3")) /
                  sum(active(d[k], p).size for k in ("cu1", "cu2", "cu3")))
    return {k: v / rms for k, v in d.items()}


def test_wall_shear_gradient_finite_differences_256():
    """256x129x256, Re 400, 50 steps, checkpoints every 10: <grad, d> against
    central differences (eps 1e-6) along 8 seeded divergence-free directions."""
    p, gy, ic, seed, amp = workloads.workload("cfg5_256x129x256")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    g = prepared(gpu_solver(p, gy), st)
    s0 = g.get_state(STATE)
    print("cfg5 gradient: 50 steps, checkpoints every 10", flush=True)
    run = WallShearRun(g, 50, checkpoint_every=10)
    L, grad = run.gradient(s0)
    print("loss", L, flush=True)
    eps = 1e-6
    worst = 0.0
    for i in range(8):
        d = divergence_free_direction(p, gy, 100 + i)
        ad = sum(np.sum((np.conj(grad[k]) * d[k]).real) for k in ("cu1", "cu2", "cu3"))
        lp = run.loss({k: s0[k] + eps * d[k] for k in STATE})
        lm = run.loss({k: s0[k] - eps * d[k] for k in STATE})
        fd = (lp - lm) / (2 * eps)
        rel = abs(ad - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"direction {i}: adjoint {ad:.12e} fd {fd:.12e} rel {rel:.2e}", flush=True)
    assert worst < 1e-6


def test_autograd_function_backward():
    """torch.autograd through WallShearLoss (16x17x16, 2 steps) equals the run's gradient."""
    p = workloads.channel_params(16, 17, 16, 200.0, ri=0.1, dt=2e-3)
    gy = workloads.channel_grid(p["ny"])
    s = seeded(p, gy, carry=False)
    run = WallShearRun(gpu_solver(p, gy), 2, checkpoint_every=1)
    rest = {k: s[k] for k in ("cp", "cf1", "cf2", "cf3", "cfth")}
    xs = [torch.tensor(s[k], requires_grad=True) for k in ("cu1", "cu2", "cu3", "cth")]
    L = WallShearLoss.apply(*xs, run, rest)
    L.backward()
    _, grad = run.gradient(s)
    for x, k in zip(xs, ("cu1", "cu2", "cu3", "cth")):
        np.testing.assert_allclose(x.grad.numpy(), grad[k], rtol=0, atol=1e-15 * np.abs(grad[k]).max())

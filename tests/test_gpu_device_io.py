"""Fields through caller device memory (dgpu_set/get_field_device and the
adjoint versions) on torch tensors and streams, and WallShearLoss on device
tensors: the same numbers as the host exchange, bit for bit."""
import numpy as np
import pytest
import torch

from diablo3_test_amd import workloads
from diablo3_test_amd.adjoint import WallShearLoss, WallShearRun
from helpers import active, gpu_solver, prepared

pytestmark = pytest.mark.gpu

SPEC = ("cu1", "cu2", "cu3", "cp", "cth", "cf1", "cf2", "cf3", "cfth")


def case(n_th=2, les=False):
    p = workloads.channel_params(32, 33, 32, 500.0, ri=0.1, dt=1e-3, les=les, n_th=n_th)
    gy = workloads.channel_grid(33)
    st = workloads.initial_state(p, gy, "vonkarman" if les else "white", 5, 1e-1 if les else 1e-2)
    return p, gy, st


def test_device_roundtrip_matches_host_exchange():
    p, gy, st = case()
    print("device field round trip 32x33x32, 2 scalars", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    g.step(2)  # a non-zero RK carry
    side = torch.cuda.Stream()
    for k in SPEC:
        host = g.get_field(k)
        with torch.cuda.stream(side):
            dev = g.get_field_device(k)
        side.synchronize()
        np.testing.assert_array_equal(dev.cpu().numpy(), host)
        # a changed field in through device memory, back out through the host
        new = dev * (1.5 + 0.25j)
        g.set_field_device(k, new)
        back = g.get_field(k)
        np.testing.assert_array_equal(active(back, p), active(new.cpu().numpy(), p))
        g.set_field(k, host)


def test_device_exchange_of_eddy_viscosity():
    p, gy, st = case(n_th=1, les=True)
    print("device nu_t / kappa_t round trip", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    g.step(1)
    for k in ("nu_t", "kappa_t"):
        host = g.get_field(k)
        dev = g.get_field_device(k)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev.cpu().numpy(), host)
        g.set_field_device(k, 2.0 * dev)
        back = g.get_field(k)
        np.testing.assert_array_equal(back[..., : p["nz"], : p["nx"]], 2.0 * host[..., : p["nz"], : p["nx"]])


def test_steps_after_device_set_match_host_set():
    """The copy on a caller stream is ordered before the handle's substeps."""
    p, gy, st = case()
    print("steps after device set vs host set", flush=True)
    a = prepared(gpu_solver(p, gy), st)
    ref = a.get_state(SPEC)
    b = gpu_solver(p, gy)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        for k in SPEC:
            b.set_field_device(k, torch.from_numpy(ref[k]).to("cuda", non_blocking=False))
    b.step(3)
    a.step(3)
    sa, sb = a.get_state(SPEC), b.get_state(SPEC)
    for k in SPEC:
        np.testing.assert_array_equal(active(sb[k], p), active(sa[k], p))


def test_adjoint_device_roundtrip():
    p, gy, st = case(n_th=1)
    g = prepared(gpu_solver(p, gy), st)
    rng = np.random.default_rng(3)
    for k in SPEC:
        shape = g.get_field(k).shape
        lam = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        g.adj_set_field_device(k, torch.from_numpy(lam).cuda())
        np.testing.assert_array_equal(active(g.adj_get_field(k), p), active(lam, p))
        dev = g.adj_get_field_device(k)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dev.cpu().numpy(), g.adj_get_field(k))


def test_wall_shear_loss_on_device_tensors():
    """torch.autograd through WallShearLoss with device tensors: the loss and
    gradients equal the host-tensor path's, and stay on the device."""
    p = workloads.channel_params(16, 17, 16, 200.0, ri=0.1, dt=2e-3)
    gy = workloads.channel_grid(p["ny"])
    st = workloads.initial_state(p, gy, "white", 4, 1e-2)
    ref_solver = prepared(gpu_solver(p, gy), st)
    s = ref_solver.get_state(SPEC)
    rest = {k: s[k] for k in ("cp", "cf1", "cf2", "cf3", "cfth")}
    names = ("cu1", "cu2", "cu3", "cth")
    # host tensors
    run_h = WallShearRun(gpu_solver(p, gy), 3, checkpoint_every=2)
    xh = [torch.tensor(s[k], requires_grad=True) for k in names]
    Lh = WallShearLoss.apply(*xh, run_h, rest)
    Lh.backward()
    # device tensors, the held fields on the device too
    run_d = WallShearRun(gpu_solver(p, gy), 3, checkpoint_every=2)
    xd = [torch.tensor(s[k], device="cuda", requires_grad=True) for k in names]
    rest_d = {k: torch.from_numpy(v).cuda() for k, v in rest.items()}
    Ld = WallShearLoss.apply(*xd, run_d, rest_d)
    Ld.backward()
    assert Ld.is_cuda and Ld.detach().item() == Lh.detach().item()
    for x, y in zip(xd, xh):
        assert x.grad.is_cuda
        np.testing.assert_array_equal(x.grad.cpu().numpy(), y.grad.numpy())

"""Flow-file transforms on the device: dgpu_get_physical / dgpu_set_physical
against the oracle's fft_xz_to_physical / fft_xz_to_fourier, and flow files
written from the device read back by the oracle (and the reverse)."""
import numpy as np
import pytest

from diablo3_test_amd import flowio, workloads
from diablo3_test_amd.binding import DgpuError
from helpers import gpu_solver, oracle_solver, rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("cu1", "cu2", "cu3", "cp", "cth")


def _case(nx, ny, nz, nth, seed, precision="fp64"):
    p = workloads.channel_params(nx, ny, nz, re=300.0, ri=0.5, dt=1e-3, n_th=nth)
    gy = workloads.channel_grid(ny)
    o = oracle_solver(p, gy)
    rng = np.random.default_rng(seed)
    shape = (ny + 2, nz + 2, nx + 2)
    for k in FIELDS:
        lead = (nth,) if k == "cth" else ()
        if k == "cth" and nth == 0:
            continue
        f = np.zeros(lead + shape)
        f[..., :nz, :nx] = rng.standard_normal(lead + (ny + 2, nz, nx))
        o.set_physical(k, f)
    g = gpu_solver(p, gy, precision)
    g.set_state(o.get_state(FIELDS if nth else FIELDS[:4]))
    return p, gy, o, g


@pytest.mark.parametrize("nx,ny,nz,nth,precision", [
    (32, 33, 32, 1, "fp64"), (64, 17, 16, 2, "fp64"), (128, 65, 64, 1, "mixed"), (16, 9, 32, 0, "fp64")])
def test_get_physical_matches_oracle(nx, ny, nz, nth, precision):
    print(f"get_physical {nx}x{ny}x{nz} nth={nth} {precision}", flush=True)
    p, gy, o, g = _case(nx, ny, nz, nth, 11, precision)
    for k in FIELDS if nth else FIELDS[:4]:
        a = g.get_physical(k)[..., :nz, :nx]
        b = o.get_physical(k)[..., :nz, :nx]
        assert rel_l2(a, b) < 1e-13, k
    g.close()


@pytest.mark.parametrize("nx,ny,nz,nth", [(32, 33, 32, 1), (256, 17, 16, 1)])
def test_set_physical_matches_oracle(nx, ny, nz, nth):
    print(f"set_physical {nx}x{ny}x{nz}", flush=True)
    p, gy, o, g = _case(nx, ny, nz, nth, 12)
    rng = np.random.default_rng(13)
    shape = (ny + 2, nz + 2, nx + 2)
    for k in FIELDS:
        lead = (nth,) if k == "cth" else ()
        f = np.zeros(lead + shape)
        f[..., :nz, :nx] = rng.standard_normal(lead + (ny + 2, nz, nx))
        g.set_physical(k, f)
        o.set_physical(k, f)
        assert rel_l2(g.get_field(k), o.get_field(k)) < 1e-13, k
    g.close()


def test_physical_rejects_other_fields():
    p, gy, o, g = _case(16, 9, 16, 0, 1)
    with pytest.raises(DgpuError):
        g.get_physical("cth")  # no scalars on this handle
    with pytest.raises(ValueError):
        g.get_physical("cf1")
    g.close()


def test_flow_file_device_to_oracle(tmp_path):
    """end.h5 written from the device, read by the oracle, and the reverse."""
    print("flow file 64x33x32", flush=True)
    p, gy, o, g = _case(64, 33, 32, 1, 14)
    f = str(tmp_path / "end.h5")
    flowio.write_flow(g, f, 2.5, 100, 1.0, 2.0, 3.0, save_pressure=True)
    flowio.write_flow(o, str(tmp_path / "ref.h5"), 2.5, 100, 1.0, 2.0, 3.0, save_pressure=True)
    from diablo3_test_amd import h5lite

    a = h5lite.read(f)["Timestep"]
    b = h5lite.read(str(tmp_path / "ref.h5"))["Timestep"]
    for k in ("U", "V", "W", "TH1", "P"):
        assert rel_l2(a[k], b[k]) < 1e-13, k
    o2 = oracle_solver(p, gy)
    info = flowio.read_flow(o2, f)
    assert info["time_step"] == 100 and info["read_pressure"]
    g2 = gpu_solver(p, gy)
    flowio.read_flow(g2, str(tmp_path / "ref.h5"))
    for k in FIELDS:
        assert rel_l2(g2.get_field(k), o2.get_field(k)) < 1e-13, k
    # the state the writers leave behind agrees too
    for k in FIELDS:
        assert rel_l2(g.get_field(k), o.get_field(k)) < 1e-13, k
    g.close()
    g2.close()


def test_flow_file_restart_matches_oracle(tmp_path):
    """A run restarted from a device-written out.h5 on the device and on the
    oracle (read_flow, the RK carry handed over, pre_first_step) stays at
    parity over two steps."""
    print("restart 32x33x32", flush=True)
    p, gy, ic, seed, amp = workloads.workload("small_32x33x32")
    g = gpu_solver(p, gy)
    g.set_state(workloads.initial_state(p, gy, ic, seed, amp))
    g.pre_first_step()
    g.step(3)
    f = str(tmp_path / "out.h5")
    flowio.write_flow(g, f, g.time, g.time_step, save_pressure=True)
    carry = {k: g.get_field(k) for k in ("cf1", "cf2", "cf3", "cfth")}
    runs = []
    for s in (gpu_solver(p, gy), oracle_solver(p, gy)):
        info = flowio.read_flow(s, f)
        s.time, s.time_step = info["time"], info["time_step"]
        s.set_state(carry)
        s.pre_first_step()
        s.step(2)
        runs.append(s)
    for k in ("cu1", "cu2", "cu3", "cp", "cth"):
        assert rel_l2(runs[0].get_field(k), runs[1].get_field(k)) < 1e-10, k
    g.close()
    runs[0].close()

"""The mixed-precision form (FP32 transforms, products and LES; FP64 state,
solves and projection) against the FP64 oracle and the FP64 GPU path.

Stated errors per field after N steps: velocity, scalar, pressure, the
discrete divergence, and the wall shear stress of statistics.f90:143-146.
The required bar for the wall shear is 1e-5 relative; the test bars
below for the other fields are the measured errors with a margin, so a
regression shows."""
import numpy as np
import pytest

from diablo3_test_amd import stats, workloads
from diablo3_test_amd.grid import metrics
from helpers import active, field_errors, gpu_solver, oracle_solver, prepared, report

pytestmark = pytest.mark.gpu

FIELDS = ("cu1", "cu2", "cu3", "cth", "cp")


def divergence_ratio(state, p, gy):
    """max |discrete divergence| on rows 2..Ny-1 over the largest term it sums."""
    _, dy, dyf = metrics(gy)
    kx, kz = workloads.wavenumbers(p)
    cu1, cu2, cu3 = (active(state[k], p) for k in ("cu1", "cu2", "cu3"))
    t1 = 1j * kx[None, None, :] * cu1
    t3 = 1j * kz[None, :, None] * cu3
    t2 = (cu2[1:] - cu2[:-1]) / dyf[:-1, None, None]
    d = (t1[:-1] + t3[:-1] + t2)[2: p["ny"]]
    scale = max(np.abs(t1).max(), np.abs(t3).max(), np.abs(t2).max())
    return np.abs(d).max() / scale


def compare(tag, mixed, ref, p, gy):
    sm, sr = mixed.get_state(FIELDS), ref.get_state(FIELDS)
    errs = field_errors(sm, sr, p, names=FIELDS)
    nu = 1.0 / p["re"]
    wm = np.array(stats.wall_shear(sm["cu1"], gy, nu))
    wr = np.array(stats.wall_shear(sr["cu1"], gy, nu))
    errs["wall_shear"] = float(np.max(np.abs(wm - wr) / np.abs(wr)))
    errs["div_mixed"] = divergence_ratio(sm, p, gy)
    report(tag, errs)
    return errs


def test_mixed_small_vs_oracle():
    p, gy, ic, seed, amp = workloads.workload("small_32x33x32")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    print("mixed 32x33x32 vs oracle, 10 steps", flush=True)
    m = prepared(gpu_solver(p, gy, precision="mixed"), st)
    o = prepared(oracle_solver(p, gy), st)
    m.step(10)
    o.step(10)
    e = compare("mixed small", m, o, p, gy)
    assert e["wall_shear"] < 1e-5
    assert max(e[k] for k in ("cu1", "cu2", "cu3", "cth")) < 1e-5
    assert e["div_mixed"] < 1e-12


def test_mixed_cfg1_vs_oracle():
    p, gy, ic, seed, amp = workloads.workload("cfg1_dns_64x65x64")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    print("mixed 64x65x64 vs oracle, 10 steps", flush=True)
    m = prepared(gpu_solver(p, gy, precision="mixed"), st)
    o = prepared(oracle_solver(p, gy), st)
    m.step(10)
    o.step(10)
    e = compare("mixed cfg1", m, o, p, gy)
    assert e["wall_shear"] < 1e-5
    assert max(e[k] for k in ("cu1", "cu2", "cu3", "cth")) < 1e-5
    assert e["div_mixed"] < 1e-12


def test_mixed_512_fifty_steps_vs_fp64():
    """512x257x512, 50 steps: mixed against the FP64 GPU path (which matches
    the oracle to the round-off floor)."""
    p, gy, ic, seed, amp = workloads.workload("dns_512x257x512")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    print("mixed vs fp64 512x257x512, 50 steps", flush=True)
    m = prepared(gpu_solver(p, gy, precision="mixed"), st)
    m.step(50)
    sm = m.get_state(FIELDS)
    m.close()
    r = prepared(gpu_solver(p, gy), st)
    r.step(50)

    class Frozen:
        def get_state(self, names):
            return {k: sm[k] for k in names}

    e = compare("mixed 512 50 steps", Frozen(), r, p, gy)
    assert e["wall_shear"] < 1e-5
    assert max(e[k] for k in ("cu1", "cu2", "cu3", "cth")) < 1e-5
    assert e["div_mixed"] < 1e-12


def test_mixed_les_vs_fp64():
    """AMD-Verstappen LES at 128x257x128, 10 steps: mixed against FP64."""
    p = workloads.channel_params(128, 257, 128, 2000.0, ri=0.1, dt=1e-3, les=True, les_model_type=5)
    gy = workloads.channel_grid(257)
    st = workloads.initial_state(p, gy, "vonkarman", 3, 1e-1)
    print("mixed LES 128x257x128 vs fp64, 10 steps", flush=True)
    m = prepared(gpu_solver(p, gy, precision="mixed"), st)
    r = prepared(gpu_solver(p, gy), st)
    m.step(10)
    r.step(10)
    e = compare("mixed les", m, r, p, gy)
    nt_m, nt_r = m.get_field("nu_t"), r.get_field("nu_t")
    e_nt = np.linalg.norm(nt_m - nt_r) / np.linalg.norm(nt_r)
    print("nu_t", f"{e_nt:.2e}", flush=True)
    assert e["wall_shear"] < 1e-5
    assert max(e[k] for k in ("cu1", "cu2", "cu3", "cth")) < 1e-5
    assert e_nt < 1e-3


@pytest.mark.parametrize("shape", [(1024, 17, 32), (32, 17, 1024), (2048, 9, 16)])
def test_long_transforms_fp64_vs_oracle(shape):
    """Transforms longer than one wavefront's 512 points (two or more
    wavefronts per transform, workgroup barriers): one step against the oracle."""
    nx, ny, nz = shape
    p = workloads.channel_params(nx, ny, nz, 400.0, ri=0.1, dt=5e-4)
    gy = workloads.channel_grid(ny)
    st = workloads.initial_state(p, gy, "white", 9, 1e-2)
    print(f"fp64 {nx}x{ny}x{nz} vs oracle, 1 step", flush=True)
    g = prepared(gpu_solver(p, gy), st)
    o = prepared(oracle_solver(p, gy), st)
    g.step(1)
    o.step(1)
    errs = field_errors(g.get_state(), o.get_state(), p)
    report(f"{nx}x{ny}x{nz}", errs)
    assert max(errs[k] for k in ("cu1", "cu2", "cu3", "cth")) < 1e-12
    assert errs["cp"] < 1e-9

"""Pins the CPU oracle (oracle/*.cpp) before it is trusted as the parity reference.

Each check is chosen so that a plausible slip in the restatement (a wrong
sign, scale, pack order or index, a dropped term) fails one of them:
closed-form single-mode transforms, numpy/pocketfft transforms, dense
solves of the Thomas systems, the divergence-free invariant of the
projection, the exact cubic amplification of Wray RK3 on a linear mode,
and discrete steady states of the mean flow.
"""
import json
import math
import os

import numpy as np
import pytest

from diablo3_test_amd import workloads
from diablo3_test_amd.grid import create_grid, metrics
from helpers import active, dptr, oracle_lib, oracle_solver, rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def small_params(nx=16, ny=9, nz=16, **kw):
    p = workloads.channel_params(nx, ny, nz, re=100.0, ri=0.0, dt=1e-3)
    p.update(kw)
    return p


def fft_fwd(s, p, f):
    """oracle fft_xz_to_fourier on a physical (Ny+2, Nz, Nx) array."""
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    buf = np.zeros((ny + 2, nz + 2, nx + 2))
    buf[:, :nz, :nx] = f
    assert oracle_lib().orc_fft_xz_to_fourier(s._h, dptr(buf)) == 0
    return buf.view(np.complex128)


def fft_inv(s, p, c):
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    buf = np.ascontiguousarray(c.astype(np.complex128)).copy()
    assert oracle_lib().orc_fft_xz_to_physical(s._h, dptr(buf.view(np.float64))) == 0
    return buf.view(np.float64).reshape(ny + 2, nz + 2, nx + 2)[:, :nz, :nx]


def grid_xz(p):
    x = np.arange(p["nx"]) * p["lx"] / p["nx"]
    z = np.arange(p["nz"]) * p["lz"] / p["nz"]
    return np.meshgrid(x, z)  # (Nz, Nx)


# ------------------------------------------------------------------ FFT ----
@pytest.mark.parametrize("nz", [16, 65])
@pytest.mark.parametrize("m,n", [(0, 0), (1, 0), (0, 2), (3, 1), (2, 5)])
def test_fft_single_mode_closed_form(m, n, nz):
    """cos(2 pi m x/Lx) cos(2 pi n z/Lz): packed coefficients 1/4 (1/2, 1) at
    (m, +-n), from the 1/(Nx Nz) scaling of fft.f90:262, 304, 310."""
    p = small_params(nz=nz)
    s = oracle_solver(p, workloads.channel_grid(p["ny"]))
    X, Z = grid_xz(p)
    f = np.cos(2 * np.pi * m * X / p["lx"]) * np.cos(2 * np.pi * n * Z / p["lz"])
    c = fft_fwd(s, p, np.broadcast_to(f, (p["ny"] + 2,) + f.shape))
    nkz = p["nz"] // 3
    exp = np.zeros((2 * nkz + 1, p["nx"] // 3 + 1), dtype=complex)
    amp = (0.5 if m else 1.0) * (0.5 if n else 1.0)
    kpos = n
    kneg = 2 * nkz + 1 - n if n else 0
    exp[kpos, m] += amp
    if n:
        exp[kneg, m] += amp
    got = active(c, p)
    for j in range(p["ny"] + 2):
        np.testing.assert_allclose(got[j], exp, atol=1e-14)


def test_fft_sign_convention():
    """sin(2 pi x/Lx) -> -i/2 at kx = 1 (forward exponent -i, fft.f90:185-204)."""
    p = small_params()
    s = oracle_solver(p, workloads.channel_grid(p["ny"]))
    X, Z = grid_xz(p)
    c = fft_fwd(s, p, np.broadcast_to(np.sin(2 * np.pi * X / p["lx"]), (p["ny"] + 2, p["nz"], p["nx"])))
    assert abs(c[3, 0, 1] - (-0.5j)) < 1e-14
    Zs = np.sin(2 * np.pi * Z / p["lz"])
    c = fft_fwd(s, p, np.broadcast_to(Zs, (p["ny"] + 2, p["nz"], p["nx"])))
    nkz = p["nz"] // 3
    assert abs(c[3, 1, 0] - (-0.5j)) < 1e-14
    assert abs(c[3, 2 * nkz, 0] - 0.5j) < 1e-14  # kz = -1 packed at 2Nkz


def test_fft_dealias_drops_high_modes():
    """Modes beyond Nkx = Nx/3, Nkz = Nz/3 vanish (fft.f90:264-266, 301-313)."""
    p = small_params()
    s = oracle_solver(p, workloads.channel_grid(p["ny"]))
    X, Z = grid_xz(p)
    nkx, nkz = p["nx"] // 3, p["nz"] // 3
    f = np.cos(2 * np.pi * (nkx + 1) * X / p["lx"]) + np.cos(2 * np.pi * (nkz + 1) * Z / p["lz"])
    c = fft_fwd(s, p, np.broadcast_to(f, (p["ny"] + 2,) + f.shape))
    assert np.abs(active(c, p)).max() < 1e-14


@pytest.mark.parametrize("nx,nz", [(16, 16), (32, 16), (16, 64), (64, 32), (16, 65), (32, 48), (16, 30),
                                   (16, 13), (16, 169)])
def test_fft_matches_numpy(nx, nz):
    """Random field: oracle == numpy (pocketfft) rfft/fft with the reference's
    scaling, dealias and kz pack; inverse == numpy irfft/ifft.  The z lengths
    that are not powers of two (65 = 5 x 13 as in the reference's front
    setups, 48, 30, the prime 13, 13 x 13) take the mixed-radix path."""
    p = small_params(nx=nx, nz=nz)
    s = oracle_solver(p, workloads.channel_grid(p["ny"]))
    rng = np.random.default_rng(0)
    f = rng.standard_normal((p["ny"] + 2, nz, nx))
    c = fft_fwd(s, p, f)
    ref = workloads.physical_to_spectral(f, nx, nz)
    np.testing.assert_allclose(active(c, p), active(ref, p), atol=1e-14)
    # inverse of a dealiased spectrum, numpy restated separately
    g = fft_inv(s, p, c)
    np.testing.assert_allclose(g, workloads.spectral_to_physical(ref, nx, nz), atol=1e-12)
    # round trip: forward(inverse(c)) == c on the dealiased block
    c2 = fft_fwd(s, p, g)
    np.testing.assert_allclose(active(c2, p), active(ref, p), atol=1e-13)


def test_c2r_ignores_imaginary_part_of_kx0():
    """FFTW c2r semantics: Im of the kx = 0 input does not reach the output."""
    p = small_params()
    s = oracle_solver(p, workloads.channel_grid(p["ny"]))
    c = np.zeros((p["ny"] + 2, p["nz"] + 2, p["nx"] // 2 + 1), dtype=complex)
    c[:, 0, 0] = 1.0 + 5.0j
    g = fft_inv(s, p, c)
    np.testing.assert_allclose(g, 1.0, atol=1e-14)


# --------------------------------------------------------------- Thomas ----
def test_thomas_real_matches_dense_solve():
    """thomas_{forward,backward}_real_mpi (channel.f90:1335-1405, 1460-1496) at
    one y-rank: rows 0..Nyp+1 of each lane."""
    rng = np.random.default_rng(1)
    nyp, inx = 7, 5
    n = nyp + 2
    a = rng.uniform(-1, 1, (n, inx))
    c = rng.uniform(-1, 1, (n, inx))
    b = 3.0 + rng.uniform(0, 1, (n, inx))
    g = rng.standard_normal((n, inx))
    x = np.zeros_like(g)
    for i in range(inx):
        A = np.diag(b[:, i]) + np.diag(a[1:, i], -1) + np.diag(c[:-1, i], 1)
        x[:, i] = np.linalg.solve(A, g[:, i])
    aa, bb, cc, gg = (np.ascontiguousarray(v.copy()) for v in (a, b, c, g))
    oracle_lib().orc_thomas_real(dptr(aa), dptr(bb), dptr(cc), dptr(gg), nyp, inx)
    np.testing.assert_allclose(gg, x, rtol=1e-13, atol=1e-14)


def test_thomas_complex_matches_dense_solve():
    """thomas_{forward,backward}_complex_mpi (channel.f90:1408-1457, 1499-1540):
    rows 1..Nyp only; rows 0 and Nyp+1 are left alone."""
    rng = np.random.default_rng(2)
    nyp, inx = 8, 4
    n = nyp + 2
    a = rng.uniform(-1, 1, (n, inx))
    c = rng.uniform(-1, 1, (n, inx))
    b = 3.0 + rng.uniform(0, 1, (n, inx))
    g = rng.standard_normal((n, inx)) + 1j * rng.standard_normal((n, inx))
    x = g.copy()
    for i in range(inx):
        sl = slice(1, nyp + 1)
        A = np.diag(b[sl, i]) + np.diag(a[2:nyp + 1, i], -1) + np.diag(c[1:nyp, i], 1)
        x[sl, i] = np.linalg.solve(A, g[sl, i])
    aa, bb, cc = (np.ascontiguousarray(v.copy()) for v in (a, b, c))
    gg = np.ascontiguousarray(g.copy())
    oracle_lib().orc_thomas_complex(dptr(aa), dptr(bb), dptr(cc), dptr(gg.view(np.float64)), nyp, inx)
    np.testing.assert_allclose(gg, x, rtol=1e-13, atol=1e-14)


# ------------------------------------------------------------ physics ----
def divergence(state, p, gy):
    """ikx cu1 + (cu2(j+1) - cu2(j))/dyf(j) + ikz cu3 on rows 2..Ny-1 (channel.f90:982-984)."""
    _, dy, dyf = metrics(gy)
    kx, kz = workloads.wavenumbers(p)
    cu1, cu2, cu3 = (active(state[k], p) for k in ("cu1", "cu2", "cu3"))
    d = 1j * kx[None, None, :] * cu1 + 1j * kz[None, :, None] * cu3
    d[:-1] += (cu2[1:] - cu2[:-1]) / dyf[:-1, None, None]
    ny = p["ny"]
    return d[2:ny]


def test_rem_div_makes_divergence_free():
    p = small_params(nx=16, ny=17, nz=16)
    gy = workloads.channel_grid(p["ny"])
    st = workloads.initial_state(p, gy, "white", seed=3, amp=0.1)
    s = oracle_solver(p, gy)
    s.set_state(st)
    d0 = np.abs(divergence(s.get_state(), p, gy)).max()
    s.rem_div_chan()
    d1 = np.abs(divergence(s.get_state(), p, gy)).max()
    assert d0 > 1e-2
    assert d1 < 1e-12 * max(1.0, d0)


def test_rk3_single_mode_amplification():
    """A y-uniform u1 = cos(kz z) under Neumann walls only feels the explicit
    horizontal viscosity: one step multiplies it by 1 + z + z^2/2 + z^3/6,
    z = -nu kz^2 dt (Wray RK3 coefficients of flow.f90:78-86)."""
    p = small_params(nx=16, ny=9, nz=16, n_th=0)
    p.update(u_bc_ymin=1, u_bc_ymax=1, px0=0.0, re=10.0, delta_t=0.05, pr=[], th_bc_ymin=[],
             th_bc_ymax=[], th_bc_ymin_c1=[], th_bc_ymax_c1=[], dthdx=[], dthdz=[])
    gy = workloads.channel_grid(p["ny"])
    s = oracle_solver(p, gy)
    shp = (p["ny"] + 2, p["nz"] + 2, p["nx"] // 2 + 1)
    cu1 = np.zeros(shp, dtype=complex)
    n = 2
    cu1[:, n, 0] = 0.5
    cu1[:, 2 * (p["nz"] // 3) + 1 - n, 0] = 0.5
    s.set_field("cu1", cu1)
    s.step(1)
    kz = workloads.wavenumbers(p)[1][n]
    z = -(1.0 / p["re"]) * kz ** 2 * p["delta_t"]
    amp = 1 + z + z * z / 2 + z ** 3 / 6
    got = s.get_field("cu1")
    np.testing.assert_allclose(got[1:p["ny"] + 1, n, 0].real, 0.5 * amp, rtol=1e-13)
    assert abs(amp - math.exp(z)) > 1e-9  # the test distinguishes RK3 from exact decay


def test_poiseuille_fixed_point_uniform_grid():
    """Uniform grid: the 3-point operator is exact on 1 - y^2, so u1 = 1 - y^2
    with px0 = -2 nu is a discrete fixed point (channel.f90:108-112, 494-505)."""
    ny = 17
    p = small_params(nx=16, ny=ny, nz=16, n_th=0, pr=[], th_bc_ymin=[], th_bc_ymax=[],
                     th_bc_ymin_c1=[], th_bc_ymax_c1=[], dthdx=[], dthdz=[])
    gy, _ = create_grid(ny, 2.0, kind=2)
    gyf, _, _ = metrics(gy)
    s = oracle_solver(p, gy)
    shp = (ny + 2, p["nz"] + 2, p["nx"] // 2 + 1)
    cu1 = np.zeros(shp, dtype=complex)
    cu1[:, 0, 0] = 1 - gyf ** 2
    s.set_field("cu1", cu1)
    s.step(3)
    got = s.get_field("cu1")[:, 0, 0].real
    np.testing.assert_allclose(got[1:ny + 1], (1 - gyf ** 2)[1:ny + 1], atol=1e-13)


def test_stretched_grid_discrete_steady_state():
    """Stretched grid: the discrete steady solution of nu D2 u = px0 with
    u = 0 at GYF(1), GYF(Ny) (solved densely here from dy, dyf) stays fixed."""
    ny = 17
    p = small_params(nx=16, ny=ny, nz=16, n_th=0, pr=[], th_bc_ymin=[], th_bc_ymax=[],
                     th_bc_ymin_c1=[], th_bc_ymax_c1=[], dthdx=[], dthdz=[])
    nu = 1.0 / p["re"]
    gy = workloads.channel_grid(ny)
    _, dy, dyf = metrics(gy)
    # rows 2..Ny-1: ((u(j+1)-u(j))/dy(j+1) - (u(j)-u(j-1))/dy(j))/dyf(j) = px0/nu
    n = ny + 2
    A = np.zeros((n, n))
    rhs = np.zeros(n)
    for j in range(n):
        if 2 <= j <= ny - 1:
            A[j, j - 1] = 1 / (dy[j] * dyf[j])
            A[j, j + 1] = 1 / (dy[j + 1] * dyf[j])
            A[j, j] = -A[j, j - 1] - A[j, j + 1]
            rhs[j] = p["px0"] / nu
        else:
            A[j, j] = 1.0
    u = np.linalg.solve(A, rhs)
    s = oracle_solver(p, gy)
    cu1 = np.zeros((n, p["nz"] + 2, p["nx"] // 2 + 1), dtype=complex)
    cu1[:, 0, 0] = u
    s.set_field("cu1", cu1)
    s.step(2)
    got = s.get_field("cu1")[:, 0, 0].real
    np.testing.assert_allclose(got[1:ny + 1], u[1:ny + 1], atol=1e-13)
    assert u.max() > 0.5  # a non-trivial profile


def test_grid_generator_matches_reference_grid_file():
    """py/setup/create_grid.py restated == test_case/grid.h5 (uniform, H = 20, N = 193)."""
    ref = json.load(open(os.path.join(GOLDEN, "test_case_grid_y.json")))
    y = np.array(ref["y"])
    G, GF = create_grid(193, 20.0, kind=3)
    np.testing.assert_allclose(G, y, rtol=0, atol=1e-12)
    gyf, dy, dyf = metrics(y)
    np.testing.assert_allclose(gyf[1], -20.0, atol=1e-12)
    np.testing.assert_allclose(gyf[193], 0.0, atol=1e-12)
    np.testing.assert_allclose(dyf[1:194], 20.0 / 192, rtol=1e-12)


def test_oracle_metrics_match_python():
    p = small_params(nx=16, ny=17, nz=16)
    gy = workloads.channel_grid(p["ny"])
    s = oracle_solver(p, gy)
    n = p["ny"] + 2
    gyf, dy, dyf = (np.zeros(n) for _ in range(3))
    kx = np.zeros(p["nx"] // 2)
    kz = np.zeros(2 * (p["nz"] // 3) + 1)
    oracle_lib().orc_metrics(s._h, dptr(gyf), dptr(dy), dptr(dyf), dptr(kx), dptr(kz))
    g2, d2, f2 = metrics(gy)
    np.testing.assert_allclose(gyf, g2, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(dy[1:], d2[1:], rtol=1e-14)
    np.testing.assert_allclose(dyf, f2, rtol=1e-14)
    kx2, kz2 = workloads.wavenumbers(p)
    np.testing.assert_array_equal(kx[: kx2.size], kx2)
    np.testing.assert_array_equal(kz, kz2)
    # pi rounded to single precision (fft.f90:152): kx(1) = 2*3.14159274.../Lx
    assert kx[1] == 2.0 * float(np.float32(np.pi)) / p["lx"]


def test_wall_shear_and_bulk_of_poiseuille():
    """stats.py on u1 = 1 - y^2 sampled at GYF: tau = 2 nu on both walls to the
    first-order one-sided difference, bulk velocity 2/3 to the trapezoid rule."""
    from diablo3_test_amd import stats
    ny = 65
    p = small_params(nx=16, ny=ny, nz=16)
    gy = workloads.channel_grid(ny)
    gyf, _, _ = metrics(gy)
    cu1 = np.zeros((ny + 2, p["nz"] + 2, p["nx"] // 2 + 1), dtype=complex)
    cu1[:, 0, 0] = 1 - gyf ** 2
    nu = 1.0 / p["re"]
    lo, hi = stats.wall_shear(cu1, gy, nu)
    h1 = gyf[2] - gyf[1]
    # (U(2) - U(1)) / h with U = 1 - y^2 at y = -1, -1 + h: exactly 2 - h
    np.testing.assert_allclose(lo, nu * (2 - h1), rtol=1e-12)
    np.testing.assert_allclose(hi, lo, rtol=1e-12)  # symmetric grid
    ub = stats.bulk_velocity(cu1, gy)
    assert abs(ub - 2.0 / 3.0) < 2e-3

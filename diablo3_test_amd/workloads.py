"""Named channel-flow workloads and their seeded initial states.

Each workload fixes the run parameters of one case of the benchmark plan:
no-slip walls, Lx = 4 pi, Lz = 2 pi, y in [-1, 1] on the tanh grid of
py/setup/grid_generators.py:56-59 (CS = 1.75, H = 2), beta = 1,
nu_v_scale = 1, no rotation, a constant mean pressure gradient
px0 = -2 nu (f_type 1, for/channel.f90:108-112) that holds a laminar
Poiseuille profile with centreline velocity 1, and one scalar with
theta = y on the walls, Pr = 1, buoyancy GRAV = (0, Ri, 0)
(for/channel.f90:162-182).

Initial states are built directly in the reference's spectral layout
(complex128, shape (Ny+2, Nz+2, Nx/2+1)); the caller projects them with the
solver's own apply_bc_vel_mpi_post + rem_div_chan (the pre_first_step
sequence of for/advance.f90:17-44).
"""
import math

import numpy as np

from .grid import create_grid, metrics


WORKLOADS = {
    # name: (Nx, Ny, Nz, Re, Ri, les, ic, seed, amplitude, dt)
    "cfg1_dns_64x65x64": (64, 65, 64, 180.0, 0.0, False, "white", 1, 1e-2, 1e-3),
    "cfg2_dns_256x193x256": (256, 193, 256, 400.0, 0.1, False, "powerlaw", 2, 5e-2, 1e-3),
    "dns_512x257x512": (512, 257, 512, 2000.0, 0.1, False, "vonkarman", 3, 1e-1, 1e-3),
    "cfg3_les_512x257x512": (512, 257, 512, 2000.0, 0.1, True, "vonkarman", 3, 1e-1, 1e-3),
    "cfg4b_1024x385x1024": (1024, 385, 1024, 2000.0, 0.1, False, "vonkarman", 3, 1e-1, 1e-3),
    "cfg5_256x129x256": (256, 129, 256, 400.0, 0.0, False, "powerlaw", 5, 5e-2, 1e-3),
    "small_32x33x32": (32, 33, 32, 180.0, 0.1, False, "white", 7, 1e-2, 1e-3),
}


def channel_params(nx, ny, nz, re, ri=0.0, dt=1e-3, les=False, les_model_type=5, n_th=1):
    """Parameter dict for ``binding.Params`` (see the module docstring)."""
    nu = 1.0 / re
    return dict(
        nx=nx, ny=ny, nz=nz, n_th=n_th,
        lx=4.0 * math.pi, lz=2.0 * math.pi,
        re=re, nu_v_scale=1.0, beta=1.0,
        ro_inv=0.0, delta=1.0,
        grav_x=0.0, grav_y=ri, grav_z=0.0,
        delta_t=dt,
        f_type=1, les_model_type=les_model_type,
        px0=-2.0 * nu, omega0=0.0, amp_omega0=0.0, force_start=0.0,
        u_bc_ymin=0, v_bc_ymin=0, w_bc_ymin=0,
        u_bc_ymax=0, v_bc_ymax=0, w_bc_ymax=0,
        u_bc_ymin_c1=0.0, v_bc_ymin_c1=0.0, w_bc_ymin_c1=0.0,
        u_bc_ymax_c1=0.0, v_bc_ymax_c1=0.0, w_bc_ymax_c1=0.0,
        pr=[1.0] * n_th,
        th_bc_ymin=[0] * n_th, th_bc_ymax=[0] * n_th,
        th_bc_ymin_c1=[-1.0] * n_th, th_bc_ymax_c1=[1.0] * n_th,
        dthdx=[0.0] * n_th, dthdz=[0.0] * n_th,
        use_les=1 if les else 0,
        create_new_flow=0,
    )


def channel_grid(ny):
    """GY(1..Ny+1) of the tanh grid, CS = 1.75, H = 2."""
    gy, _ = create_grid(ny, 2.0, kind=5, CS=1.75)
    return gy


def workload(name):
    """(params dict, gy, ic kind, seed, amplitude) of a named workload."""
    nx, ny, nz, re, ri, les, ic, seed, amp, dt = WORKLOADS[name]
    return channel_params(nx, ny, nz, re, ri, dt, les), channel_grid(ny), ic, seed, amp


def wavenumbers(p):
    """kx(0..Nx/3), kz(0..2Nkz) exactly as for/fft.f90:152, 213-230 (single-precision pi)."""
    pi = float(np.float32(4.0) * np.arctan(np.float32(1.0)))
    nkx, nkz = p["nx"] // 3, p["nz"] // 3
    kx = np.arange(nkx + 1) * (2.0 * pi) / p["lx"]
    kz = np.zeros(2 * nkz + 1)
    kz[: nkz + 1] = np.arange(nkz + 1) * (2.0 * pi) / p["lz"]
    for k in range(1, nkz + 1):
        kz[2 * nkz + 1 - k] = -k * (2.0 * pi) / p["lz"]
    return kx, kz


def physical_to_spectral(f, nx, nz):
    """Real (Ny+2, Nz, Nx) -> packed spectral (Ny+2, Nz+2, Nx/2+1), the forward
    transform of fft_xz_to_fourier (for/fft.f90:247-315) done with numpy."""
    nkx, nkz = nx // 3, nz // 3
    F = np.fft.rfft(f, axis=2) / nx
    F[:, :, nkx + 1:] = 0.0
    F = np.fft.fft(F, axis=1) / nz
    out = np.zeros((f.shape[0], nz + 2, nx // 2 + 1), dtype=np.complex128)
    out[:, : nkz + 1, : nkx + 1] = F[:, : nkz + 1, : nkx + 1]
    out[:, nkz + 1: 2 * nkz + 1, : nkx + 1] = F[:, nz - nkz: nz, : nkx + 1]
    return out


def spectral_to_physical(c, nx, nz):
    """Inverse of the above (fft_xz_to_physical, for/fft.f90:319-382)."""
    nkx, nkz = nx // 3, nz // 3
    ny2 = c.shape[0]
    F = np.zeros((ny2, nz, nx // 2 + 1), dtype=np.complex128)
    F[:, : nkz + 1, : nkx + 1] = c[:, : nkz + 1, : nkx + 1]
    F[:, nz - nkz: nz, : nkx + 1] = c[:, nkz + 1: 2 * nkz + 1, : nkx + 1]
    F = np.fft.ifft(F, axis=1) * nz
    F[:, :, 0] = F[:, :, 0].real  # c2r reads only the real part at kx = 0
    return np.fft.irfft(F, n=nx, axis=2) * nx


def _random_spectrum(rng, p, kx, kz, env_y, shape_fn, amp):
    """Random packed spectrum with a given radial shape, Hermitian at kx = 0."""
    nx, nz, ny = p["nx"], p["nz"], p["ny"]
    nkx, nkz = nx // 3, nz // 3
    K = np.sqrt(kx[None, :] ** 2 + kz[:, None] ** 2)  # (2nkz+1, nkx+1)
    shape = np.where(K > 0, shape_fn(np.where(K > 0, K, 1.0)), 0.0)
    c = np.zeros((ny + 2, nz + 2, nx // 2 + 1), dtype=np.complex128)
    z = rng.standard_normal((ny + 2, 2 * nkz + 1, nkx + 1)) + 1j * rng.standard_normal(
        (ny + 2, 2 * nkz + 1, nkx + 1)
    )
    blk = z * shape[None]
    # Hermitian symmetry of the kx = 0 column: c(-kz) = conj(c(kz)), c(0,0) real
    for k in range(1, nkz + 1):
        blk[:, 2 * nkz + 1 - k, 0] = np.conj(blk[:, k, 0])
    blk[:, 0, 0] = 0.0
    # physical rms of the perturbation at an envelope maximum -> amp
    w = np.ones(nkx + 1)
    w[1:] = 2.0
    power = (np.abs(blk) ** 2 * w[None, None, :]).sum(axis=(1, 2)).mean()
    blk *= amp / math.sqrt(max(power, 1e-300))
    c[:, : 2 * nkz + 1, : nkx + 1] = blk * env_y[:, None, None]
    return c


def initial_state(p, gy, ic, seed, amp):
    """Seeded initial state: Poiseuille u1 = 1 - y^2, theta = y, cp = 0, plus
    perturbations on u1, u2, u3 (``white``: i.i.d. N(0, amp^2) at interior
    points in physical space; ``powerlaw``: random-phase modes with energy
    ~ k^-2; ``vonkarman``: E(k) ~ k^4 / (1 + (k/k0)^2)^(17/6)), each with a
    (1 - y^2) envelope in the spectral cases.  cf1..3 and cfth start at zero."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    gyf, _, _ = metrics(gy)
    GY = np.zeros(ny + 2)
    GY[1:] = gy
    shp = (ny + 2, nz + 2, nx // 2 + 1)
    cu = [np.zeros(shp, dtype=np.complex128) for _ in range(3)]
    cu[0][:, 0, 0] = 1.0 - gyf ** 2
    if ic == "white":
        for f in range(3):
            noise = np.zeros((ny + 2, nz, nx))
            noise[2:ny, :, :] = rng.normal(0.0, amp, size=(ny - 2, nz, nx))
            cu[f] += physical_to_spectral(noise, nx, nz)
    else:
        kx, kz = wavenumbers(p)
        if ic == "powerlaw":
            fn = lambda K: 1.0 / K  # energy ~ k^-2
        elif ic == "vonkarman":
            k0 = 4.0
            fn = lambda K: np.sqrt(K ** 4 / (1.0 + (K / k0) ** 2) ** (17.0 / 6.0) / K)
        else:
            raise ValueError(ic)
        for f in range(3):
            yy = GY if f == 1 else gyf
            env = np.clip(1.0 - yy ** 2, 0.0, None)
            cu[f] += _random_spectrum(rng, p, kx, kz, env, fn, amp)
    state = {"cu1": cu[0], "cu2": cu[1], "cu3": cu[2], "cp": np.zeros(shp, dtype=np.complex128)}
    n_th = p["n_th"]
    if n_th:
        cth = np.zeros((n_th,) + shp, dtype=np.complex128)
        cth[:, :, 0, 0] = gyf[None, :]
        state["cth"] = cth
    return state

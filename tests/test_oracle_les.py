"""The oracle's subgrid models (les_chan, les_chan_th of for/les.f90) against
closed forms: the eddy viscosity and diffusivity from their tensor
definitions (les_forms.py), the explicit subgrid forcing as the divergence
of the modelled stress, the discrete steady state of a Smagorinsky channel,
and the start-up gate of for/channel.f90:259."""
import numpy as np
import pytest
from scipy.optimize import fsolve

from diablo3_test_amd import workloads
from diablo3_test_amd.grid import metrics
from helpers import active, oracle_solver, prepared, rel_l2
import les_forms

TYPES = (1, 4, 5)


def les_params(typ, nx=32, ny=33, nz=32, **kw):
    p = workloads.channel_params(nx, ny, nz, 2000.0, ri=0.1, dt=1e-3, les=True, les_model_type=typ)
    p.update(kw)
    return p


def bc_state(p, gy, seed=3, amp=1e-1):
    """A seeded turbulent-like state after pre_first_step and
    apply_BC_vel_mpi_post, with the scalar wall planes as
    apply_BC_th_{lower,upper}_post leaves them for Dirichlet walls
    (for/boundary.f90:728-821), so that the wall updates at the start of
    les_chan change nothing and a DNS substep sees the same state."""
    st = workloads.initial_state(p, gy, "vonkarman", seed, amp)
    # scalar fluctuations of the velocity's spectrum shape on top of the mean profile
    fluct = st["cu1"][2: p["ny"]].copy()
    fluct[:, 0, 0] = 0.0
    for n in range(p["n_th"]):
        st["cth"][n, 2: p["ny"]] += 0.5 * fluct
    o = prepared(oracle_solver(p, gy), st)
    o.apply_bc_vel_mpi_post()
    s = o.get_state()
    ny = p["ny"]
    for n in range(p["n_th"]):
        for j, c1 in ((0, p["th_bc_ymin_c1"][n]), (1, p["th_bc_ymin_c1"][n]),
                      (ny, p["th_bc_ymax_c1"][n]), (ny + 1, p["th_bc_ymax_c1"][n])):
            s["cth"][n, j] = 0.0
            s["cth"][n, j, 0, 0] = c1
    return s


def physical(s, name, p):
    return s.get_field(name)[..., : p["nz"], : p["nx"]]


@pytest.mark.parametrize("typ", TYPES)
def test_eddy_viscosity_closed_form(typ):
    p = les_params(typ)
    gy = workloads.channel_grid(p["ny"])
    st = bc_state(p, gy)
    o = oracle_solver(p, gy)
    o.set_state(st)
    o.rk_chan_1(1, 0.0, 1)
    F = les_forms.Fields(st, p, gy)
    want = les_forms.nu_t(F, typ)
    got = physical(o, "nu_t", p)
    scale = np.abs(want).max()
    assert scale > 1e-4
    # the AMD models switch off where the numerator is negative: a real share of points stays on
    assert np.count_nonzero(want) > 0.1 * want[3: p["ny"]].size
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * scale)
    kt = physical(o, "kappa_t", p)[0]
    if typ in (1, 4):
        np.testing.assert_array_equal(kt[1:], got[1:])  # kappa_t = nu_t (les.f90:1142-1150)
    else:
        kw = np.zeros_like(kt)
        for j in range(3, p["ny"]):
            kw[j] = les_forms.kappa_amd(F, 0, j, at_gy=True)
        ks = np.abs(kw).max()
        assert ks > 1e-5
        np.testing.assert_allclose(kt, kw, rtol=0, atol=1e-11 * ks)


@pytest.mark.parametrize("typ", TYPES)
def test_explicit_forcing_is_stress_divergence(typ):
    """The subgrid terms one substep adds to cf1..3 and cfth: LES run minus a
    DNS run from the same state (the remaining explicit terms are identical)."""
    p = les_params(typ)
    gy = workloads.channel_grid(p["ny"])
    st = bc_state(p, gy)
    names = ("cf1", "cf2", "cf3", "cfth")
    out = {}
    for les in (1, 0):
        q = dict(p, use_les=les)
        o = oracle_solver(q, gy)
        o.set_state(st)
        o.rk_chan_1(1, 0.0, 1)
        out[les] = o.get_state(names)
    F = les_forms.Fields(st, p, gy)
    ny, js = p["ny"], 2  # no-slip walls: jstart = 2
    f1, f2, f3 = les_forms.velocity_forcing(F, typ, js)
    nt = les_forms.nu_t(F, typ)
    fth = les_forms.scalar_forcing(F, typ, 0, 2, ny - 1, nt)
    rows = {"cf1": slice(js, ny), "cf3": slice(js, ny), "cf2": slice(2, ny + 1), "cfth": slice(2, ny)}
    want = {"cf1": f1, "cf2": f2, "cf3": f3, "cfth": fth}
    for k in names:
        d = out[1][k] - out[0][k]
        if k == "cfth":
            d = d[0]
        r = rows[k]
        got_k, want_k = active(d[r], p), active(want[k][r], p)
        err = rel_l2(got_k, want_k)
        print(typ, k, f"|f| = {np.linalg.norm(want_k):.3e}", f"rel err = {err:.2e}")
        assert np.linalg.norm(want_k) > 1e-6
        assert err < 1e-9, (k, err)
        # outside those rows the substep adds nothing
        mask = np.ones(d.shape[0], bool)
        mask[r] = False
        assert np.abs(active(d[mask], p)).max() <= 1e-12 * np.abs(active(out[0][k], p)).max()


def smagorinsky_steady_state(p, gy):
    """Discrete steady mean flow and scalar of a Smagorinsky channel.  With
    U(y) only, the subgrid stress reduces to nu_t(j) = (C_s D_j)^2 |dU/dy|
    on GY points 3..Ny-1 (D_j^3 = 9 dx dz 2 dy(j)), and the steady momentum
    balance is the flux form
        ((nu + nu_t) dU/dy |_(j+1) - (nu + nu_t) dU/dy |_j) / dyf(j) = px0
    on rows 2..Ny-1 with U = 0 on GYF points 0, 1, Ny, Ny+1; the scalar
    balance is the same with nu/Pr + nu_t and no source."""
    _, dy, dyf = metrics(gy)
    ny = p["ny"]
    nu = 1.0 / p["re"]
    dxs, dzs = 3.0 * p["lx"] / p["nx"], 3.0 * p["lz"] / p["nz"]
    c2d2 = np.zeros(ny + 2)
    for j in range(3, ny):
        c2d2[j] = les_forms.C_SMAG ** 2 * (dxs * 2.0 * dy[j] * dzs) ** (2.0 / 3.0)

    def full(x, lo, hi):
        u = np.zeros(ny + 2)
        u[:2], u[ny:] = lo, hi
        u[2:ny] = x
        return u

    def eddy(u):
        nt = np.zeros(ny + 2)
        for j in range(3, ny):
            nt[j] = c2d2[j] * abs(u[j] - u[j - 1]) / dy[j]
        return nt

    def residual(x, diff, nt_of, src, lo, hi):
        u = full(x, lo, hi)
        nt = nt_of(u)
        flux = [(diff + nt[j]) * (u[j] - u[j - 1]) / dy[j] for j in range(ny + 1)] + [0.0]
        return np.array([(flux[j + 1] - flux[j]) / dyf[j] - src for j in range(2, ny)])

    y = metrics(gy)[0]
    x0 = 1 - y[2:ny] ** 2
    xu = fsolve(residual, x0, args=(nu, eddy, p["px0"], 0.0, 0.0), xtol=1e-15)
    u = full(xu, 0.0, 0.0)
    assert np.abs(residual(xu, nu, eddy, p["px0"], 0.0, 0.0)).max() < 1e-12 * abs(p["px0"]) * 1e3
    nt = eddy(u)
    lo, hi = p["th_bc_ymin_c1"][0], p["th_bc_ymax_c1"][0]
    xt = fsolve(residual, y[2:ny], args=(nu / p["pr"][0], lambda _: nt, 0.0, lo, hi), xtol=1e-15)
    th = full(xt, lo, hi)
    return u, th, nt


@pytest.mark.parametrize("typ", TYPES)
def test_mean_shear_steady_state(typ):
    """Smagorinsky: the laminar-like state with the eddy viscosity of its own
    shear stays fixed over two RK3 steps (implicit and explicit halves of the
    eddy-viscosity CN terms, channel.f90:519-547, 676-689, 792-805).  The AMD
    models vanish on a pure shear, so there the DNS steady state holds and
    nu_t stays exactly zero."""
    p = les_params(typ, nx=16, ny=17, nz=16)
    p.update(re=100.0, px0=-2.0 / 100.0, grav_y=0.0)
    gy = workloads.channel_grid(p["ny"])
    if typ == 1:
        u, th, nt = smagorinsky_steady_state(p, gy)
        assert nt.max() > 1.0 / p["re"]  # the eddy viscosity dominates
    else:
        u, th, _ = dns_steady(p, gy)
    ny = p["ny"]
    o = oracle_solver(p, gy)
    shp = (ny + 2, p["nz"] + 2, p["nx"] // 2 + 1)
    cu1 = np.zeros(shp, dtype=complex)
    cu1[:, 0, 0] = u
    cth = np.zeros((1,) + shp, dtype=complex)
    cth[0, :, 0, 0] = th
    o.set_field("cu1", cu1)
    o.set_field("cth", cth)
    o.step(2)
    got_u = o.get_field("cu1")[:, 0, 0].real
    got_t = o.get_field("cth")[0, :, 0, 0].real
    np.testing.assert_allclose(got_u[1: ny + 1], u[1: ny + 1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(got_t[1: ny + 1], th[1: ny + 1], rtol=0, atol=1e-12)
    got_nt = physical(o, "nu_t", p)
    if typ == 1:
        np.testing.assert_allclose(got_nt[:, 0, 0], nt, rtol=1e-12, atol=1e-15)
    else:
        assert not got_nt.any()


def dns_steady(p, gy):
    """Steady mean flow and scalar without an eddy viscosity."""
    _, dy, dyf = metrics(gy)
    ny = p["ny"]
    nu = 1.0 / p["re"]
    n = ny + 2

    def solve(diff, src, lo, hi):
        A = np.zeros((n, n))
        rhs = np.zeros(n)
        for j in range(n):
            if 2 <= j <= ny - 1:
                A[j, j - 1] = diff / (dy[j] * dyf[j])
                A[j, j + 1] = diff / (dy[j + 1] * dyf[j])
                A[j, j] = -A[j, j - 1] - A[j, j + 1]
                rhs[j] = src
            else:
                A[j, j] = 1.0
                rhs[j] = lo if j <= 1 else hi
        return np.linalg.solve(A, rhs)

    u = solve(nu, p["px0"], 0.0, 0.0)
    th = solve(nu / p["pr"][0], 0.0, p["th_bc_ymin_c1"][0], p["th_bc_ymax_c1"][0])
    return u, th, None


def test_startup_gate():
    """With create_new_flow the model stays off for the first 100 steps
    (channel.f90:259): the substep then equals a DNS substep exactly (nu_t is
    still zero, so the eddy-viscosity CN terms add exact zeros)."""
    p = les_params(5, create_new_flow=1)
    gy = workloads.channel_grid(p["ny"])
    st = bc_state(p, gy)
    res = {}
    for les, ts in ((1, 100), (0, 100), (1, 101)):
        o = oracle_solver(dict(p, use_les=les), gy)
        o.set_state(st)
        o.rk_chan_1(1, 0.0, ts)
        res[(les, ts)] = o.get_state(("cu1", "cu2", "cu3", "cth", "cf1"))
    for k in res[(0, 100)]:
        np.testing.assert_array_equal(res[(1, 100)][k], res[(0, 100)][k])
    assert rel_l2(active(res[(1, 101)]["cf1"], p), active(res[(0, 100)]["cf1"], p)) > 1e-6


def test_rejects_unknown_model_type():
    p = les_params(2)
    gy = workloads.channel_grid(p["ny"])
    with pytest.raises(Exception):
        oracle_solver(p, gy)

"""FP64 PyTorch restatement of the LES stage of the channel substep, for autograd.

It restates, at one y-rank, what rk_chan_1 runs when the LES gate is open
(for/channel.f90:259-289): the post boundary conditions of the state
(for/boundary.f90:299-498, 728-821), apply_BC_les / apply_BC_th_les
(for/les.f90:2001-2100), les_chan for the Smagorinsky (type 1, :94-270),
AMD-Rozema (type 4, :279-565) and AMD-Verstappen (type 5, :570-850) models
with the stress divergence and explicit cross terms (:857-919), and
les_chan_th (:1137-1415).  The eddy-viscosity parts of the CN solves live in
TorchChannel.cn_solve.

Plane ranges follow the reference loop by loop; values the reference leaves
stale in scratch planes that no later loop reads are zero here.  Branches
whose derivative is singular where they are not taken (the AMD clip's
division, Smagorinsky's square root) are masked on both sides, so autograd
differentiates the branch the forward took without producing NaN from the
other.

Arrays: spectral (planes, 2*Nkz+1, Nkx+1) complex128, physical
(planes, Nz, Nx) float64, planes 0..Ny+1.
"""
import math

import torch

C_SMAG, C_AMD, BETA_SGS = 0.17, 0.2887, 3.0  # les.f90:41-60


def _set(a, idx, val):
    """a with a[idx] = val (autograd-friendly: writes into a clone)."""
    b = a.clone()
    b[idx] = val
    return b


def bc_vel_post(tc, cu1, cu2, cu3, time):
    """apply_BC_vel_mpi_post on the spectral velocity (boundary.f90:299-498)."""
    p, ny, dy, dyf = tc.p, tc.ny, tc.dy, tc.dyf
    z = torch.zeros_like(cu1[0])

    def lower(c, bc, c1):
        if bc == 0:
            c = _set(c, 1, z)
            c = _set(c, 0, z)
            c = _set(c, (1, 0, 0), c1)
            return _set(c, (0, 0, 0), c1)
        c = _set(c, 1, c[2])
        c = _set(c, 0, c[1])
        c = _set(c, (1, 0, 0), c[2, 0, 0] - dy[2] * c1)
        return _set(c, (0, 0, 0), c[1, 0, 0] - dy[1] * c1)

    def upper(c, bc, c1, top_ghost_zero=False):
        if bc == 0:
            c = _set(c, ny, z)
            c = _set(c, ny + 1, z)
            c = _set(c, (ny, 0, 0), c1)
            c = _set(c, (ny + 1, 0, 0), c1)
            if top_ghost_zero:  # "Ghost cell not used" (boundary.f90:442)
                c = _set(c, (ny + 1, 0, 0), 0.0)
            return c
        c = _set(c, ny, c[ny - 1])
        c = _set(c, ny + 1, c[ny])
        c = _set(c, (ny, 0, 0), c[ny - 1, 0, 0] + dy[ny] * c1)
        return _set(c, (ny + 1, 0, 0), c[ny, 0, 0] + dy[ny] * c1)

    cu1 = lower(cu1, p["u_bc_ymin"], p["u_bc_ymin_c1"])
    cu3 = lower(cu3, p["w_bc_ymin"], p["w_bc_ymin_c1"])
    if p["v_bc_ymin"] == 0:
        cu2 = _set(cu2, 1, 2.0 * p["v_bc_ymin_c1"] - cu2[2])
        cu2 = _set(cu2, 0, cu2[1])
    else:
        cu2 = _set(cu2, 1, cu2[2])
        cu2 = _set(cu2, 0, cu2[1])
        cu2 = _set(cu2, (1, 0, 0), cu2[2, 0, 0] - dyf[1] * p["v_bc_ymin_c1"])
        cu2 = _set(cu2, (0, 0, 0), cu2[1, 0, 0] - dyf[1] * p["v_bc_ymin_c1"])
    cu1 = upper(cu1, p["u_bc_ymax"], p["u_bc_ymax_c1"])
    wc1 = p["w_bc_ymax_c1"]
    if p["w_bc_ymax"] != 0 and p["f_type"] == 4:
        wc1 = p["w_bc_ymax_c1"] + p["amp_omega0"] * math.sin(
            p["omega0"] * (p["ro_inv"] / p["delta"] * time - p["force_start"]))
        tc.w_transient = wc1
    cu3 = upper(cu3, p["w_bc_ymax"], wc1, top_ghost_zero=True)
    if p["v_bc_ymax"] == 0:
        # only the (0,0) mode is written (boundary.f90:476-480)
        cu2 = _set(cu2, (ny + 1, 0, 0), 2.0 * p["v_bc_ymax_c1"] - cu2[ny, 0, 0])
    else:
        cu2 = _set(cu2, ny, cu2[ny - 1])
        cu2 = _set(cu2, ny + 1, cu2[ny])
        cu2 = _set(cu2, (ny + 1, 0, 0), cu2[ny, 0, 0] + dy[ny] * p["v_bc_ymax_c1"])
    return cu1, cu2, cu3


def bc_th_post(tc, cth):
    """apply_BC_th_{lower,upper}_post for every scalar (boundary.f90:728-821)."""
    p, ny, dy = tc.p, tc.ny, tc.dy
    out = []
    for n in range(tc.nth):
        c = cth[n]
        z = torch.zeros_like(c[0])
        c1l, c1u = p["th_bc_ymin_c1"][n], p["th_bc_ymax_c1"][n]
        if p["th_bc_ymin"][n] == 0:
            c = _set(_set(c, 0, z), 1, z)
            c = _set(_set(c, (1, 0, 0), c1l), (0, 0, 0), c1l)
        else:
            c = _set(c, 1, c[2])
            c = _set(c, 0, c[1])
            c = _set(c, (1, 0, 0), c[2, 0, 0] - dy[2] * c1l)
            c = _set(c, (0, 0, 0), c[1, 0, 0] - dy[1] * c1l)
        if p["th_bc_ymax"][n] == 0:
            c = _set(_set(c, ny, z), ny + 1, z)
            c = _set(_set(c, (ny, 0, 0), c1u), (ny + 1, 0, 0), c1u)
        else:
            c = _set(c, ny, c[ny - 1])
            c = _set(c, ny + 1, c[ny])
            c = _set(c, (ny, 0, 0), c[ny - 1, 0, 0] + dy[ny] * c1u)
            c = _set(c, (ny + 1, 0, 0), c[ny, 0, 0] + dy[ny] * c1u)
        out.append(c)
    return torch.stack(out) if out else cth


def bc_les(tc, cu1, cu3, cth):
    """apply_BC_les + apply_BC_th_les: Neumann walls copy the interior into the
    ghost planes, every mode (les.f90:2001-2100)."""
    p, ny = tc.p, tc.ny

    def up(c):
        c = _set(c, ny, c[ny - 1])
        return _set(c, ny + 1, c[ny])

    def dn(c):
        c = _set(c, 1, c[2])
        return _set(c, 0, c[1])

    if p["u_bc_ymax"] == 1:
        cu1 = up(cu1)
    if p["w_bc_ymax"] == 1:
        cu3 = up(cu3)
    if p["u_bc_ymin"] == 1:
        cu1 = dn(cu1)
    if p["w_bc_ymin"] == 1:
        cu3 = dn(cu3)
    ths = []
    for n in range(tc.nth):
        c = cth[n]
        if p["th_bc_ymax"][n] == 1:
            c = up(c)
        if p["th_bc_ymin"][n] == 1:
            c = dn(c)
        ths.append(c)
    return cu1, cu3, (torch.stack(ths) if ths else cth)


class _Planes:
    """Geometry helpers shared by the model evaluations."""

    def __init__(self, tc):
        self.tc = tc
        self.ny = tc.ny
        self.n = tc.ny + 2
        self.dy = tc.dy
        self.dyf = tc.dyf

    def col(self, v):
        return v[:, None, None]

    def mask(self, lo, hi):
        m = torch.zeros((self.n, 1, 1), dtype=torch.float64)
        m[lo: hi + 1] = 1.0
        return m

    def gy_interp(self, a):
        """GYF -> GY at planes 1..Ny+1 (les.f90:150-165): (a(j) dyf(j-1) + a(j-1) dyf(j)) / (2 dy(j))."""
        dy, dyf = self.dy, self.dyf
        out = torch.zeros_like(a)
        out[1:] = (a[1:] * self.col(dyf[:-1]) + a[:-1] * self.col(dyf[1:])) / (2.0 * self.col(dy[1:]))
        return out

    def avg_up(self, a):
        """GY -> GYF at planes 0..Ny: 0.5 (a(j+1) + a(j))."""
        out = torch.zeros_like(a)
        out[:-1] = 0.5 * (a[1:] + a[:-1])
        return out

    def ddy_dy(self, a):
        """(a(j) - a(j-1)) / dy(j) at planes 1..Ny+1, zero at plane 0."""
        out = torch.zeros_like(a)
        out[1:] = (a[1:] - a[:-1]) / self.col(self.dy[1:])
        return out


def _spectral_strain(tc, cu1, cu2, cu3, invariant):
    """compute_strain_chan / _invariant (les.f90:1463-1577), spectral."""
    ny, dy, dyf = tc.ny, tc.dy, tc.dyf
    p = tc.p
    dxs = (p["lx"] / tc.nx) * BETA_SGS
    dzs = (p["lz"] / tc.nz) * BETA_SGS
    ikx, ikz = tc.ikx, tc.ikz
    c0 = ikx * cu1
    c2 = ikz * cu3
    if invariant:
        c4 = 0.5 * (ikz * cu1 * (dzs / dxs) + ikx * cu3 * (dxs / dzs))
    else:
        c4 = 0.5 * (ikz * cu1 + ikx * cu3)
    c1 = torch.zeros_like(cu2)
    c1[: ny + 1] = (cu2[1:] - cu2[:-1]) / dyf[: ny + 1, None, None]
    c1[ny + 1] = -10000.0  # ghost_CS2_mpi's sentinel (les.f90:2532)
    c3 = torch.zeros_like(cu1)
    c5 = torch.zeros_like(cu1)
    d1 = (cu1[1:] - cu1[:-1]) / dy[1:, None, None]
    d3 = (cu3[1:] - cu3[:-1]) / dy[1:, None, None]
    if invariant:
        dys = (dy[1:] * 2.0)[:, None, None]
        c3[1:] = 0.5 * (d1 * (dys / dxs) + ikx * cu2[1:] * (dxs / dys))
        c5[1:] = 0.5 * (ikz * cu2[1:] * (dzs / dys) + d3 * (dys / dzs))
    else:
        c3[1:] = 0.5 * (d1 + ikx * cu2[1:])
        c5[1:] = 0.5 * (ikz * cu2[1:] + d3)
    # SIJ1..6 order: 11, 22, 33, 12, 13, 23
    return [c0, c1, c2, c3, c4, c5]


def _spectral_rotation(tc, cu1, cu2, cu3, invariant):
    """compute_rotation_chan / _invariant (les.f90:1584-1675): O12, O13, O23."""
    dy = tc.dy
    p = tc.p
    dxs = (p["lx"] / tc.nx) * BETA_SGS
    dzs = (p["lz"] / tc.nz) * BETA_SGS
    ikx, ikz = tc.ikx, tc.ikz
    if invariant:
        o5 = 0.5 * (ikz * cu1 * (dzs / dxs) - ikx * cu3 * (dxs / dzs))
    else:
        o5 = 0.5 * (ikz * cu1 - ikx * cu3)
    o4 = torch.zeros_like(cu1)
    o6 = torch.zeros_like(cu1)
    d1 = (cu1[1:] - cu1[:-1]) / dy[1:, None, None]
    d3 = (cu3[1:] - cu3[:-1]) / dy[1:, None, None]
    if invariant:
        dys = (dy[1:] * 2.0)[:, None, None]
        o4[1:] = 0.5 * (d1 * (dys / dxs) - ikx * cu2[1:] * (dxs / dys))
        o6[1:] = 0.5 * (ikz * cu2[1:] * (dzs / dys) - d3 * (dys / dzs))
    else:
        o4[1:] = 0.5 * (d1 - ikx * cu2[1:])
        o6[1:] = 0.5 * (ikz * cu2[1:] - d3)
    return o4, o5, o6


def _clip_div(v, den):
    """max(0, v) / den where v > 0, else 0 (les.f90's AMD clip), masked on both sides."""
    pos = v > 0.0
    return torch.where(pos, v, torch.zeros_like(v)) / torch.where(pos, den, torch.ones_like(den))


def _amd_rozema(a, w, dxq, dyq, dzq):
    """The AMD-Rozema point model of les.f90:306-361 (weights per direction)."""
    a1, a2, a3, a4, a5, a6 = a
    w4, w5, w6 = w
    v = (dxq * a1 ** 3 + dyq * a2 ** 3 + dzq * a3 ** 3 + (2 * dxq + dyq) * a1 * a4 * a4 +
         (2 * dxq + dzq) * a1 * a5 * a5 + (2 * dyq + dxq) * a2 * a4 * a4 + (2 * dyq + dzq) * a2 * a6 * a6 +
         (2 * dzq + dxq) * a3 * a5 * a5 + (2 * dzq + dyq) * a3 * a6 * a6 + 2 * (dxq + dyq + dzq) * a4 * a5 * a6)
    v = (-v - dyq * a1 * w4 * w4 - dzq * a1 * w5 * w5 - dxq * a2 * w4 * w4 - dzq * a2 * w6 * w6 -
         dxq * a3 * w5 * w5 - dyq * a3 * w6 * w6 - 2 * dzq * a4 * w5 * w6 + 2 * dyq * a5 * w4 * w6 -
         2 * dxq * a6 * w4 * w5)
    den = (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 + 2 * a6 * a6 +
           2 * w6 * w6)
    return _clip_div(v, den)


def _amd_verstappen(a, w):
    """The AMD-Verstappen point model of les.f90:593-642 (invariant tensors)."""
    a1, a2, a3, a4, a5, a6 = a
    w4, w5, w6 = w
    v = (a1 ** 3 + a2 ** 3 + a3 ** 3 + 3 * a1 * a4 * a4 + 3 * a1 * a5 * a5 + 3 * a2 * a4 * a4 +
         3 * a2 * a6 * a6 + 3 * a3 * a5 * a5 + 3 * a3 * a6 * a6 + 6 * a4 * a5 * a6)
    v = (-v - a1 * w4 * w4 - a1 * w5 * w5 - a2 * w4 * w4 - a2 * w6 * w6 - a3 * w5 * w5 - a3 * w6 * w6 -
         2 * a4 * w5 * w6 + 2 * a5 * w4 * w6 - 2 * a6 * w4 * w5)
    den = (a1 * a1 + a2 * a2 + a3 * a3 + 2 * a4 * a4 + 2 * w4 * w4 + 2 * a5 * a5 + 2 * w5 * w5 + 2 * a6 * a6 +
           2 * w6 * w6)
    return _clip_div(v, den)


def _safe_sqrt(x, m):
    """sqrt(x) on the planes where m is 1, zero elsewhere (no gradient there)."""
    on = (m > 0).expand_as(x)
    return torch.where(on, torch.sqrt(torch.where(on, x, torch.ones_like(x))), torch.zeros_like(x))


def les_chan(tc, cu1, cu2, cu3, u1, u3):
    """les_chan (les.f90:20-919) on the post-BC spectral velocity cu*, with u1, u3
    its physical form.  Returns (nu_t, df1, df2, df3): the eddy viscosity in
    physical space and the spectral increments of cf1, cf2, cf3."""
    p, ny = tc.p, tc.ny
    pl = _Planes(tc)
    typ = p["les_model_type"]
    dy, dyf = tc.dy, tc.dyf
    dx1, dz1 = p["lx"] / tc.nx, p["lz"] / tc.nz
    P = tc.to_phys
    js = tc.js
    j1, j2 = js, ny - 1  # the last-rank branch of les.f90:63-66
    S = [P(c) for c in _spectral_strain(tc, cu1, cu2, cu3, False)]
    m_gyf = pl.mask(1, ny)
    du1 = pl.ddy_dy(u1)
    du3 = pl.ddy_dy(u3)
    delta_y = torch.zeros(ny + 2, dtype=torch.float64)
    delta_yf = torch.zeros(ny + 2, dtype=torch.float64)
    if typ == 1:
        gi = [pl.gy_interp(x) for x in S]
        s4u, s6u = pl.avg_up(S[3]), pl.avg_up(S[5])
        ss = _safe_sqrt(2.0 * S[0] ** 2 + 4.0 * s4u ** 2 + 4.0 * S[4] ** 2 + 2.0 * S[1] ** 2 + 4.0 * s6u ** 2 +
                        2.0 * S[2] ** 2, m_gyf)
        m_gy = pl.mask(1, ny + 1)
        tp = _safe_sqrt(2.0 * gi[0] ** 2 + 4.0 * S[3] ** 2 + 4.0 * gi[4] ** 2 + 2.0 * gi[1] ** 2 + 4.0 * S[5] ** 2 +
                        2.0 * gi[2] ** 2, m_gy)
        s1, s5, s3 = [m_gyf * x * ss + (1 - m_gyf) * x for x in (S[0], S[4], S[2])]
        s4 = m_gy * tp * (S[3] - 0.5 * du1) + (1 - m_gy) * S[3]
        s6 = m_gy * tp * (S[5] - 0.5 * du3) + (1 - m_gy) * S[5]
        for j in range(1, ny + 1):
            delta_yf[j] = -2.0 * C_SMAG ** 2 * (dx1 * BETA_SGS * float(dyf[j]) * 1.0 * dz1 * BETA_SGS) ** (2.0 / 3.0)
        for j in range(1, ny + 2):
            delta_y[j] = -2.0 * C_SMAG ** 2 * (dx1 * BETA_SGS * float(dy[j]) * 2.0 * dz1 * BETA_SGS) ** (2.0 / 3.0)
        nt = pl.mask(2, ny) * (-0.5 * pl.col(delta_y) * tp)
    elif typ == 4:
        o4, o5, o6 = [P(c) for c in _spectral_rotation(tc, cu1, cu2, cu3, False)]
        dxq, dzq = (dx1 * BETA_SGS) ** 2, (dz1 * BETA_SGS) ** 2
        dyq = pl.col((dy * 2.0) ** 2)  # les.f90:306 (dy, not dyf, at the GYF points too)
        a_f = (S[0], S[1], S[2], pl.avg_up(S[3]), S[4], pl.avg_up(S[5]))
        w_f = (pl.avg_up(o4), o5, pl.avg_up(o6))
        ss = m_gyf * _amd_rozema(a_f, w_f, dxq, dyq, dzq)
        gi = [pl.gy_interp(x) for x in S]
        a_g = (gi[0], gi[1], gi[2], S[3], gi[4], S[5])
        w_g = (o4, pl.gy_interp(o5), o6)
        m_gy = pl.mask(2, ny + 1)
        tp = m_gy * _amd_rozema(a_g, w_g, dxq, dyq, dzq)
        # S1, S5, S3 scaled on planes 1..Ny+1 (the Ny+1 plane is never read)
        s1, s5, s3 = [m_gyf * x * ss + (1 - m_gyf) * x for x in (S[0], S[4], S[2])]
        s4 = m_gy * tp * (S[3] - 0.5 * du1) + (1 - m_gy) * S[3]
        s6 = m_gy * tp * (S[5] - 0.5 * du3) + (1 - m_gy) * S[5]
        delta_y[1:] = -2.0 * C_AMD ** 2
        delta_yf[1:] = -2.0 * C_AMD ** 2
        nt = pl.mask(2, ny) * (-0.5 * pl.col(delta_y) * tp)
    elif typ == 5:
        I = [P(c) for c in _spectral_strain(tc, cu1, cu2, cu3, True)]
        o4, o5, o6 = [P(c) for c in _spectral_rotation(tc, cu1, cu2, cu3, True)]
        m_f = pl.mask(tc.js, tc.je)
        a_f = (I[0], I[1], I[2], pl.avg_up(I[3]), I[4], pl.avg_up(I[5]))
        w_f = (pl.avg_up(o4), o5, pl.avg_up(o6))
        ss = m_f * _amd_verstappen(a_f, w_f)
        gi = [pl.gy_interp(x) for x in I]
        a_g = (gi[0], gi[1], gi[2], I[3], gi[4], I[5])
        w_g = (o4, pl.gy_interp(o5), o6)
        tp = pl.mask(2, ny) * _amd_verstappen(a_g, w_g)  # TEMP at Ny+1 is zero
        s1, s5, s3 = [m_f * x * ss + (1 - m_f) * x for x in (S[0], S[4], S[2])]
        m_gy = pl.mask(2, ny + 1)
        s4 = m_gy * tp * (S[3] - 0.5 * du1) + (1 - m_gy) * S[3]
        s6 = m_gy * tp * (S[5] - 0.5 * du3) + (1 - m_gy) * S[5]
        for j in range(1, ny + 2):
            delta_yf[j] = -2.0 * C_AMD ** 2 * 3.0 / (1.0 / (dx1 * BETA_SGS) ** 2 + 1.0 / (float(dyf[j]) * 2.0) ** 2 +
                                                    1.0 / (dz1 * BETA_SGS) ** 2)
            delta_y[j] = -2.0 * C_AMD ** 2 * 3.0 / (1.0 / (dx1 * BETA_SGS) ** 2 + 1.0 / (float(dy[j]) * 2.0) ** 2 +
                                                   1.0 / (dz1 * BETA_SGS) ** 2)
        nt = pl.mask(2, ny + 1) * (-0.5 * pl.col(delta_y) * tp)
    else:
        raise ValueError("les_model_type must be 1, 4 or 5")
    # ghost_les_mpi at NprocY = 1: nu_t zero at planes 1, 2, Ny, Ny+1 (and 0, never set)
    keep = torch.ones((ny + 2, 1, 1), dtype=torch.float64)
    keep[[0, 1, 2, ny, ny + 1]] = 0.0
    nt = keep * nt
    F = tc.to_spec
    c1, c3, c4, c5, c6 = F(s1), F(s3), F(s4), F(s5), F(s6)
    mf = pl.mask(1, ny)
    c1 = mf * pl.col(delta_yf) * c1 + (1 - mf) * c1
    c5 = mf * pl.col(delta_yf) * c5 + (1 - mf) * c5
    c3 = mf * pl.col(delta_yf) * c3 + (1 - mf) * c3
    mg = pl.mask(1, ny + 1)
    c4 = mg * pl.col(delta_y) * c4 + (1 - mg) * c4
    c6 = mg * pl.col(delta_y) * c6 + (1 - mg) * c6
    ikx, ikz = tc.ikx, tc.ikz
    zero = torch.zeros_like(c4[:1])
    c4p = torch.cat([c4[1:], zero])  # c4(j+1)
    c6p = torch.cat([c6[1:], zero])
    md = pl.mask(j1, j2)
    dyfc = pl.col(dyf)
    df1 = -md * (ikx * c1 + (c4p - c4) / dyfc + ikz * c5)
    df3 = -md * (ikx * c5 + (c6p - c6) / dyfc + ikz * c3)
    m2 = pl.mask(2, ny)
    df2 = -m2 * (ikx * c4 + ikz * c6)
    # explicit cross terms (les.f90:890-919)
    df2 = df2 + m2 * ikx * F(m2 * nt * du1) + m2 * ikz * F(m2 * nt * du3)
    return nt, df1, df2, df3


def les_chan_th(tc, cth, ths, nt, grads):
    """les_chan_th (les.f90:1091-1415) on the post-BC spectral scalars cth (ths
    physical).  grads: the physical velocity gradients of
    compute_all_gradients_chan (type 5 only).  Returns (kappa_t list,
    dfth list)."""
    p, ny = tc.p, tc.ny
    pl = _Planes(tc)
    typ = p["les_model_type"]
    dy, dyf = tc.dy, tc.dyf
    dx1, dz1 = p["lx"] / tc.nx, p["lz"] / tc.nz
    P, F = tc.to_phys, tc.to_spec
    ikx, ikz = tc.ikx, tc.ikz
    kts, dfs = [], []
    for n in range(tc.nth):
        mt = pl.mask(tc.jst[n], tc.jet[n])
        c = cth[n]
        m1 = pl.mask(1, ny + 1)
        if typ in (1, 4):
            kt = m1 * nt  # constant SGS Prandtl number 1 (les.f90:1137-1150)
            df = torch.zeros_like(c)
            for ik in (ikx, ikz):
                g = P(m1 * ik * c)
                df = df + mt * ik * F(m1 * kt * g)
            kts.append(kt)
            dfs.append(df)
            continue
        d11, d22, d33, d21, d31, d13, d23, d12, d32 = grads
        tx = P(m1 * ikx * c)
        tz = P(m1 * ikz * c)
        ty = P(pl.ddy_dy(c))
        deltax, deltaz = dx1 * BETA_SGS, dz1 * BETA_SGS
        dx2, dz2 = deltax * deltax, deltaz * deltaz
        # GYF points, planes jstart_th..jend_th (les.f90:1236-1280)
        deltay = pl.col(dyf * 2.0)
        dy2 = deltay * deltay
        tyf = pl.avg_up(ty)
        v = -(dx2 * d11 * tx * tx + dy2 * d22 * tyf * tyf + dz2 * d33 * tz * tz +
              (dy2 * pl.avg_up(d12) + dx2 * pl.avg_up(d21)) * tx * tyf +
              (dz2 * d13 + dx2 * d31) * tx * tz +
              (dz2 * pl.avg_up(d23) + dy2 * pl.avg_up(d32)) * tyf * tz)
        st = mt * _clip_div(v, (deltax * tx) ** 2 + (deltay * tyf) ** 2 + (deltaz * tz) ** 2)
        # GY points, planes 2..Ny+1 (les.f90:1285-1330)
        deltay = pl.col(dy * 2.0)
        dy2 = deltay * deltay
        txg, tzg = pl.gy_interp(tx), pl.gy_interp(tz)
        v = -(dx2 * pl.gy_interp(d11) * txg * txg + dy2 * pl.gy_interp(d22) * ty * ty +
              dz2 * pl.gy_interp(d33) * tzg * tzg + (dy2 * d12 + dx2 * d21) * txg * ty +
              (dz2 * pl.gy_interp(d13) + dx2 * pl.gy_interp(d31)) * txg * tzg +
              (dz2 * d23 + dy2 * d32) * ty * tzg)
        tt = pl.mask(2, ny + 1) * _clip_div(v, (deltax * txg) ** 2 + (deltay * ty) ** 2 + (deltaz * tzg) ** 2)
        fx = mt * st * tx + (1 - mt) * tx
        fz = mt * st * tz + (1 - mt) * tz
        dlt_y = torch.zeros(ny + 2, dtype=torch.float64)
        dlt_yf = torch.zeros(ny + 2, dtype=torch.float64)
        for j in range(1, ny + 2):
            dlt_yf[j] = -2.0 * C_AMD ** 2 * 3.0 / (1.0 / (dx1 * BETA_SGS) ** 2 + 1.0 / (float(dyf[j]) * 2.0) ** 2 +
                                                  1.0 / (dz1 * BETA_SGS) ** 2)
            dlt_y[j] = -2.0 * C_AMD ** 2 * 3.0 / (1.0 / (dx1 * BETA_SGS) ** 2 + 1.0 / (float(dy[j]) * 2.0) ** 2 +
                                                 1.0 / (dz1 * BETA_SGS) ** 2)
        keep = torch.ones((ny + 2, 1, 1), dtype=torch.float64)
        keep[[0, 1, 2, ny, ny + 1]] = 0.0
        kt = keep * pl.mask(2, ny) * (-0.5 * pl.col(dlt_y) * tt)
        cx, cz = F(fx), F(fz)
        mg = pl.mask(1, ny + 1)
        cx = mg * 0.5 * pl.col(dlt_yf) * cx + (1 - mg) * cx
        cz = mg * 0.5 * pl.col(dlt_yf) * cz + (1 - mg) * cz
        kts.append(kt)
        dfs.append(-mt * (ikx * cx + ikz * cz))
    return kts, dfs


def all_gradients(tc, cu1, cu2, cu3):
    """compute_all_gradients_chan (les.f90:1681-1767), velocity part, physical:
    d11, d22, d33, d21, d31, d13, d23, d12, d32 (dij = d u_i / d x_j)."""
    dy, dyf = tc.dy, tc.dyf
    P = tc.to_phys
    ikx, ikz = tc.ikx, tc.ikz
    d12 = torch.zeros_like(cu1)
    d22 = torch.zeros_like(cu2)
    d32 = torch.zeros_like(cu3)
    d12[1:] = (cu1[1:] - cu1[:-1]) / dy[1:, None, None]
    d22[1:] = (cu2[1:] - cu2[:-1]) / dyf[:-1, None, None]  # les.f90:1713, as written there
    d32[1:] = (cu3[1:] - cu3[:-1]) / dy[1:, None, None]
    return [P(ikx * cu1), P(d22), P(ikz * cu3), P(ikx * cu2), P(ikx * cu3), P(ikz * cu1), P(ikz * cu2),
            P(d12), P(d32)]


def les_stage(tc, cu1, cu2, cu3, cth, time):
    """The LES branch of rk_chan_1 (channel.f90:259-278): post BCs of the state,
    les_chan, les_chan_th.  Returns the post-BC spectral state, its physical
    form, nu_t, kappa_t and the increments of cf1..3, cfth."""
    cu1, cu2, cu3 = bc_vel_post(tc, cu1, cu2, cu3, time)
    cth = bc_th_post(tc, cth)
    cu1, cu3, cth = bc_les(tc, cu1, cu3, cth)
    u1, u2, u3 = tc.to_phys(cu1), tc.to_phys(cu2), tc.to_phys(cu3)
    ths = [tc.to_phys(cth[n]) for n in range(tc.nth)]
    nt, df1, df2, df3 = les_chan(tc, cu1, cu2, cu3, u1, u3)
    grads = all_gradients(tc, cu1, cu2, cu3) if tc.p["les_model_type"] == 5 else None
    kts, dfth = les_chan_th(tc, cth, ths, nt, grads)
    return dict(cu1=cu1, cu2=cu2, cu3=cu3, cth=cth, u1=u1, u2=u2, u3=u3, ths=ths, nu_t=nt, kappa_t=kts,
                df1=df1, df2=df2, df3=df3, dfth=dfth)

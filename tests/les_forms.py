"""Closed forms of the subgrid models in tensor notation, evaluated with numpy.

These are the checks for the oracle's les_chan / les_chan_th: the oracle
follows for/les.f90's expanded polynomials and staggered-grid loops; here each
model is written from its definition on the velocity-gradient tensor
G[i, k] = d u_i / d x_k (rows: component, columns: direction), assembled at
the two y locations of the staggered grid:

* GY points (where u2 lives): d/dy of u1, u3 and u2 itself are native there,
  GYF quantities are interpolated with les.f90:150-165's weights;
* GYF points (where u1, u3, th live): GY quantities are averaged from the
  two neighbouring GY points (les.f90:172-182).

Model definitions (filter widths D_k):
* Smagorinsky: nu = (C_s D)^2 sqrt(2 S_ij S_ij), D^3 = D_x D_y D_z;
* AMD: nu = C^2 max(0, -(D_k^2 G_ik G_jk S_ij)) / (G_lm G_lm) (Rozema);
* AMD in scaled variables (Verstappen): G^_ik = (D_k / D_i) G_ik,
  nu = C^2 3/(sum_k D_k^-2) max(0, -G^_ik G^_jk S^_ij) / (G^_lm G^_lm);
* AMD scalar: kappa = C^2 3/(sum_k D_k^-2) max(0, -D_k^2 G_ik T_k T_i) / (D_l T_l)^2.
"""
import numpy as np

from diablo3_test_amd import workloads
from diablo3_test_amd.grid import metrics

C_SMAG = 0.17     # les.f90:41
C_AMD = 0.2887    # les.f90:42
BETA_SGS = 3.0    # les.f90:60


class Fields:
    """Physical fields and their horizontal derivatives on planes 0..Ny+1."""

    def __init__(self, state, p, gy):
        self.p = p
        nx, nz = p["nx"], p["nz"]
        self.gyf, self.dy, self.dyf = metrics(gy)
        kx, kz = workloads.wavenumbers(p)
        nkx, nkz = nx // 3, nz // 3
        shp = state["cu1"].shape
        self.KX = np.zeros(shp[1:])
        self.KZ = np.zeros(shp[1:])
        self.KX[: 2 * nkz + 1, : nkx + 1] = kx[None, :]
        self.KZ[: 2 * nkz + 1, : nkx + 1] = kz[:, None]
        phys = lambda c: workloads.spectral_to_physical(c, nx, nz)
        self.phys = phys
        self.spec = lambda f: workloads.physical_to_spectral(f, nx, nz)
        self.u = [phys(state[k]) for k in ("cu1", "cu2", "cu3")]
        self.ux = [phys(1j * self.KX * state[k]) for k in ("cu1", "cu2", "cu3")]
        self.uz = [phys(1j * self.KZ * state[k]) for k in ("cu1", "cu2", "cu3")]
        cth = state.get("cth")
        if cth is not None and cth.shape[0] > 0:
            self.th = [phys(c) for c in cth]
            self.thx = [phys(1j * self.KX * c) for c in cth]
            self.thz = [phys(1j * self.KZ * c) for c in cth]
        else:
            self.th = self.thx = self.thz = []
        self.ny = p["ny"]
        self.dxs = p["lx"] / nx * BETA_SGS
        self.dzs = p["lz"] / nz * BETA_SGS

    # d/dx and d/dz of a physical field, through the dealiased spectrum
    def ddx(self, f):
        return self.phys(1j * self.KX * self.spec(f))

    def ddz(self, f):
        return self.phys(1j * self.KZ * self.spec(f))

    def dy_gy(self, f, j):
        """(f(j) - f(j-1)) / dy(j): d/dy of a GYF field at GY point j."""
        return (f[j] - f[j - 1]) / self.dy[j]

    def dy_gyf(self, f, j):
        """(f(j+1) - f(j)) / dyf(j): d/dy of a GY field at GYF point j."""
        return (f[j + 1] - f[j]) / self.dyf[j]

    def to_gy(self, a, b, j):
        """GYF values a = f(j), b = f(j-1) interpolated to GY point j."""
        return (a * self.dyf[j - 1] + b * self.dyf[j]) / (2.0 * self.dy[j])

    def grad_gy(self, j):
        """G at GY point j (2 <= j <= Ny)."""
        u, ux, uz = self.u, self.ux, self.uz
        gi = lambda f: self.to_gy(f[j], f[j - 1], j)
        g22 = self.to_gy(self.dy_gyf(u[1], j), self.dy_gyf(u[1], j - 1), j)
        return np.array([
            [gi(ux[0]), self.dy_gy(u[0], j), gi(uz[0])],
            [ux[1][j], g22, uz[1][j]],
            [gi(ux[2]), self.dy_gy(u[2], j), gi(uz[2])],
        ])

    def grad_gyf(self, j):
        """G at GYF point j (1 <= j <= Ny)."""
        u, ux, uz = self.u, self.ux, self.uz
        av = lambda a, b: 0.5 * (a + b)
        return np.array([
            [ux[0][j], av(self.dy_gy(u[0], j + 1), self.dy_gy(u[0], j)), uz[0][j]],
            [av(ux[1][j + 1], ux[1][j]), self.dy_gyf(u[1], j), av(uz[1][j + 1], uz[1][j])],
            [ux[2][j], av(self.dy_gy(u[2], j + 1), self.dy_gy(u[2], j)), uz[2][j]],
        ])

    def scaled_gy(self, j):
        """G^_ik = (D_k / D_i) G_ik at GY point j with D = (dxs, 2 dy(j), dzs)."""
        D = np.array([self.dxs, 2.0 * self.dy[j], self.dzs])
        return self.grad_gy(j) * (D[None, :] / D[:, None])[:, :, None, None]

    def scaled_gyf(self, j):
        """G^ at GYF point j: the x-z entries scaled at j, the entries that
        live at GY points (12, 21, 23, 32) averaged from the scaled values of
        GY points j and j+1, each with its own 2 dy (les.f90:1527-1559)."""
        G = self.grad_gyf(j).copy()
        D = np.array([self.dxs, 1.0, self.dzs])
        for (i, k) in ((0, 0), (1, 1), (2, 2), (0, 2), (2, 0)):
            G[i, k] = G[i, k] * D[k] / D[i]
        a, b = self.scaled_gy(j + 1), self.scaled_gy(j)
        for (i, k) in ((0, 1), (1, 0), (1, 2), (2, 1)):
            G[i, k] = 0.5 * (a[i, k] + b[i, k])
        return G


def sym(G):
    return 0.5 * (G + G.transpose(1, 0, 2, 3))


def smag(G, dprod):
    """Smagorinsky eddy viscosity for filter-width product dprod = D_x D_y D_z."""
    S = sym(G)
    return C_SMAG ** 2 * dprod ** (2.0 / 3.0) * np.sqrt(2.0 * np.einsum("ij...,ij...->...", S, S))


def amd_rozema(G, D2):
    """AMD with filter widths D2 = (D_x^2, D_y^2, D_z^2), in the form the
    expanded polynomial of les.f90:306-361 takes: the numerator symmetrised
    over G and its transpose, -1/2 D_k^2 (G_ik G_jk + G_ki G_kj) S_ij, which
    is Rozema's -(D_k d_k u_i)(D_k d_k u_j) S_ij when the widths are equal."""
    S = sym(G)
    D2 = np.asarray(D2)
    num = -0.5 * (np.einsum("k,ik...,jk...,ij...->...", D2, G, G, S)
                  + np.einsum("k,ki...,kj...,ij...->...", D2, G, G, S))
    den = np.einsum("ij...,ij...->...", G, G)
    return C_AMD ** 2 * np.where(num > 0, num / np.where(den > 0, den, 1.0), 0.0)


def amd_verstappen(Gs, D):
    """AMD in scaled variables; D = (D_x, D_y, D_z) sets the prefactor."""
    Ss = sym(Gs)
    num = -np.einsum("ik...,jk...,ij...->...", Gs, Gs, Ss)
    den = np.einsum("ij...,ij...->...", Gs, Gs)
    pref = 3.0 / (D[0] ** -2 + D[1] ** -2 + D[2] ** -2)
    return C_AMD ** 2 * pref * np.where(num > 0, num / np.where(den > 0, den, 1.0), 0.0)


def amd_scalar(G, T, D):
    """AMD scalar diffusivity for gradient T = (T_x, T_y, T_z) and widths D."""
    D = np.asarray(D)
    num = -np.einsum("k,ik...,k...,i...->...", D ** 2, G, T, T)
    den = sum((D[k] * T[k]) ** 2 for k in range(3))
    pref = 3.0 / (D[0] ** -2 + D[1] ** -2 + D[2] ** -2)
    return C_AMD ** 2 * pref * np.where(num > 0, num / np.where(den > 0, den, 1.0), 0.0)


def nu_gy(F, typ, j):
    """Eddy viscosity from the model at GY point j, before the wall zeroing."""
    if typ == 1:
        return smag(F.grad_gy(j), F.dxs * 2.0 * F.dy[j] * F.dzs)
    if typ == 4:
        return amd_rozema(F.grad_gy(j), (F.dxs ** 2, (2.0 * F.dy[j]) ** 2, F.dzs ** 2))
    return amd_verstappen(F.scaled_gy(j), (F.dxs, 2.0 * F.dy[j], F.dzs))


def nu_gyf(F, typ, j):
    """The model's eddy viscosity at GYF point j, the factor of the GYF
    stresses S11, S13, S33.  Smagorinsky uses D_y = dyf(j) there
    (les.f90:203), Rozema-AMD D_y = 2 dy(j) (les.f90:306), Verstappen-AMD
    the prefactor with D_y = 2 dyf(j) (les.f90:775-779)."""
    if typ == 1:
        return smag(F.grad_gyf(j), F.dxs * F.dyf[j] * F.dzs)
    if typ == 4:
        return amd_rozema(F.grad_gyf(j), (F.dxs ** 2, (2.0 * F.dy[j]) ** 2, F.dzs ** 2))
    return amd_verstappen(F.scaled_gyf(j), (F.dxs, 2.0 * F.dyf[j], F.dzs))


def nu_t(F, typ):
    """nu_t on planes 0..Ny+1: the model at GY points 3..Ny-1, zero at the
    planes next to the walls (les.f90:2250-2265)."""
    out = np.zeros_like(F.u[0])
    for j in range(3, F.ny):
        out[j] = nu_gy(F, typ, j)
    return out


def kappa_amd(F, n, j, at_gy):
    """AMD scalar diffusivity of scalar n at GY (at_gy) or GYF point j.  The
    wall-normal derivative of u2 enters at the location les.f90:1713 gives
    it: du2dy(j) = (u2(j) - u2(j-1)) / dyf(j-1), one GYF point below j."""
    th, thx, thz = F.th[n], F.thx[n], F.thz[n]
    G22f = lambda jj: F.dy_gyf(F.u[1], jj)
    if at_gy:
        G = F.grad_gy(j)
        G[1, 1] = F.to_gy(G22f(j - 1), G22f(j - 2), j)
        T = np.array([F.to_gy(thx[j], thx[j - 1], j), F.dy_gy(th, j), F.to_gy(thz[j], thz[j - 1], j)])
        D = (F.dxs, 2.0 * F.dy[j], F.dzs)
    else:
        G = F.grad_gyf(j)
        G[1, 1] = G22f(j - 1)
        T = np.array([thx[j], 0.5 * (F.dy_gy(th, j + 1) + F.dy_gy(th, j)), thz[j]])
        D = (F.dxs, 2.0 * F.dyf[j], F.dzs)
    return amd_scalar(G, T, D)


def velocity_forcing(F, typ, js):
    """The explicit subgrid forcing of u1, u2, u3 (spectral, planes 0..Ny+1)
    that the substep adds to the RK terms: the x/z divergence of the model
    stress 2 nu S_ij and the y divergence of its (1,2), (3,2) parts less the
    d(u1)/dy, d(u3)/dy pieces that the implicit solve carries
    (les.f90:857-919).  The GY stresses use the model value on GY points
    2..Ny (before the wall zeroing); the cross terms use nu_t.  Rows: u1, u3
    on js..Ny-1, u2 on 2..Ny."""
    ny = F.ny
    u, ux, uz = F.u, F.ux, F.uz
    nraw = np.zeros_like(u[0])
    for j in range(2, ny + 1):
        nraw[j] = nu_gy(F, typ, j)
    nt = nu_t(F, typ)
    spec, kx, kz = F.spec, 1j * F.KX, 1j * F.KZ
    plane = lambda f: spec(f[None])[0]
    f1 = np.zeros((ny + 2,) + F.KX.shape, dtype=complex)
    f2 = np.zeros_like(f1)
    f3 = np.zeros_like(f1)
    # tau12 = -2 nu (S12 - du1/dy / 2) = -nu du2/dx, likewise tau23
    t12 = spec(-nraw * ux[1])
    t23 = spec(-nraw * uz[1])
    for j in range(js, ny):
        nf = nu_gyf(F, typ, j)
        s11, s33 = ux[0][j], uz[2][j]
        s13 = 0.5 * (uz[0][j] + ux[2][j])
        f1[j] = (kx * plane(2 * nf * s11) + kz * plane(2 * nf * s13)
                 - (t12[j + 1] - t12[j]) / F.dyf[j])
        f3[j] = (kx * plane(2 * nf * s13) + kz * plane(2 * nf * s33)
                 - (t23[j + 1] - t23[j]) / F.dyf[j])
    for j in range(2, ny + 1):
        c1 = plane(nt[j] * F.dy_gy(u[0], j))
        c3 = plane(nt[j] * F.dy_gy(u[2], j))
        f2[j] = -kx * t12[j] - kz * t23[j] + kx * c1 + kz * c3
    return f1, f2, f3


def scalar_forcing(F, typ, n, jst, jet, nt):
    """The explicit subgrid forcing of scalar n: d/dx (k d th/dx) + d/dz (k d th/dz)
    with k = nu_t at the same plane index for the constant-Prandtl closures
    (les.f90:1142-1203) and the AMD scalar diffusivity at GYF points for
    Verstappen-AMD (les.f90:1238-1275, 1392-1413)."""
    spec, kx, kz = F.spec, 1j * F.KX, 1j * F.KZ
    out = np.zeros((F.ny + 2,) + F.KX.shape, dtype=complex)
    for j in range(jst, jet + 1):
        k = nt[j] if typ in (1, 4) else kappa_amd(F, n, j, at_gy=False)
        out[j] = kx * spec((k * F.thx[n][j])[None])[0] + kz * spec((k * F.thz[n][j])[None])[0]
    return out

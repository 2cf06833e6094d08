"""Mean-flow diagnostics of the channel state: the subset of DIABLO's
statistics (for/statistics.f90) that defines the wall shear stress and the
bulk velocity, computed from the spectral state in the reference layout.

* ume(j) = Re cu1(0, 0, j)                                  (statistics.f90:728)
* dudy(j) = (ume(j) - ume(j-1)) / (gyf(j) - gyf(j-1)), j = 1..Ny  (:143-146)
* integrate_y_var: sum over j = 2..Ny of (v(j) + v(j-1))/2 dy(j), over Ly (:1974-1990)

The no-slip walls sit at GYF points 1 and Ny.  The wall shear stress on each
wall is nu times the one-sided dU/dy there: tau_lower = nu dudy(2),
tau_upper = -nu dudy(Ny) (both positive for a Poiseuille profile).
"""
import numpy as np

from .grid import metrics


def mean_profile(cu1):
    """ume(0..Ny+1): the (0, 0) mode of u1."""
    return np.ascontiguousarray(cu1[..., :, 0, 0].real)


def dudy(ume, gy):
    """dU/dy at GY points 1..Ny (index 0 unused)."""
    gyf, _, _ = metrics(gy)
    out = np.zeros_like(ume)
    out[..., 1:-1] = (ume[..., 1:-1] - ume[..., :-2]) / (gyf[1:-1] - gyf[:-2])
    return out


def wall_shear(cu1, gy, nu):
    """(tau_lower, tau_upper) of a state, or of a stack of states."""
    ny = len(gy) - 1
    d = dudy(mean_profile(cu1), gy)
    return nu * d[..., 2], -nu * d[..., ny]


def bulk_velocity(cu1, gy, ly=None):
    """integrate_y_var of ume divided by Ly (default: the wall-to-wall height)."""
    gyf, dy, _ = metrics(gy)
    ny = len(gy) - 1
    ume = mean_profile(cu1)
    if ly is None:
        ly = gyf[ny] - gyf[1]
    res = sum(0.5 * (ume[..., j] + ume[..., j - 1]) * dy[j] for j in range(2, ny + 1))
    return res / ly

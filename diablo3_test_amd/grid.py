"""Wall-normal grid of the channel: the generator of py/setup/create_grid.py.

``create_grid(N, H, kind, CS)`` returns (G, GF) exactly as
py/setup/create_grid.py:22-72 does: G holds the N+1 values that DIABLO reads
from grid.h5:/grids/y into GY(1..N+1) (for/domain.f90:429-483), GF the
fractional grid.  The stretching functions are those of
py/setup/grid_generators.py:36-64.
"""
import numpy as np


def _uniform_positive(J, CS):  # grid_generators.py:36-39 (returns J-1, as the reference does)
    return J - 1


def _uniform_symmetric(J, CS):  # :41-44
    return J - 0.5


def _uniform_negative(J, CS):  # :46-49
    return J - 1


def _ends_positive(J, CS):  # :51-54
    return (np.tanh(CS * (2 * J - 1)) / np.tanh(CS) + 1) / 2


def _ends_symmetric(J, CS):  # :56-59
    return np.tanh(CS * (2 * J - 1.0)) / np.tanh(CS) / 2


def _centre_symmetric(J, CS):  # :61-64
    return (CS * (2 * J - 1.0) ** 3 + 2 * J - 1.0) / (CS + 1) / 2


GRID_GENERATORS = {  # grid_generators.py:67-74
    1: _uniform_positive,
    2: _uniform_symmetric,
    3: _uniform_negative,
    4: _ends_positive,
    5: _ends_symmetric,
    6: _centre_symmetric,
}


def create_grid(N, H, kind=5, CS=1.75):
    """(G, GF) for N wall-normal points on a domain of height H (create_grid.py:22-72)."""
    J = np.arange(-1, N + 2, 1, dtype=np.float64) / N
    G = H * GRID_GENERATORS[kind](J, CS)
    GF = (G[:-1] + G[1:]) / 2
    gf_low, gf_up = GF[1], GF[N]
    g_low, g_up = G[1], G[N + 1]
    GF = GF * (g_up - g_low) / (gf_up - gf_low)
    G = G * (g_up - g_low) / (gf_up - gf_low)
    shift = g_low - GF[1]
    GF = GF + shift
    G = G + shift
    return G[1:-1], GF


def metrics(gy):
    """GYF, DY, DYF (index 0..N+1) from GY(1..N+1) as for/domain.f90:230-254, 501-527."""
    gy = np.asarray(gy, dtype=np.float64)
    N = gy.size - 1
    GY = np.zeros(N + 2)
    GY[1:] = gy
    gyf = np.zeros(N + 2)
    gyf[1:N + 1] = 0.5 * (GY[1:N + 1] + GY[2:N + 2])
    gyf[0] = 2.0 * gyf[1] - gyf[2]
    gyf[N + 1] = 2.0 * gyf[N] - gyf[N - 1]
    dy = np.zeros(N + 2)
    dy[1:] = gyf[1:] - gyf[:-1]
    dyf = np.zeros(N + 2)
    dyf[1:N + 1] = GY[2:N + 2] - GY[1:N + 1]
    dyf[0] = dyf[1]
    dyf[N + 1] = dyf[N]
    return gyf, dy, dyf

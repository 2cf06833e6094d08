"""DIABLO flow files (start.h5, out.NNNNNN.h5, end.h5) for a ChannelSolver.

The format is that of WriteHDF5 / ReadHDF5 (phdf5.f90:913-1459):

* file attributes  Resolution (int32[3] = Nx, Ny, Nz) and Date (20 one-byte
  strings, 'DD-MON-YYYY HH:MM:SS', phdf5.f90:1562-1618);
* group /Timestep with attributes Time, Save_Flow_Time, Save_Stats_Time,
  Save_Movie_Time (float64) and Time_Step (int32);
* datasets /Timestep/U, V, W, TH1.., and P when the pressure is saved: the
  physical fields on planes y = 1..Ny as float64 arrays of Fortran shape
  (Nx, Ny, Nz), i.e. C shape (Nz, Ny, Nx) (swapzy/swapyz, phdf5.f90:1462-1497).
  V lives on the GY points in the solver and is stored interpolated to the
  GYF points (g2gf, phdf5.f90:1543-1559); reading inverts that (gf2g,
  phdf5.f90:1500-1540).

The transforms between spectral and physical space run on the device
(ChannelSolver.get_physical / set_physical); this module does the layout
and the y interpolations on the host.  The reference writes the datasets
chunked by (Nx, 1, Nz); this writer stores them contiguously, which HDF5
readers treat alike.  The reader accepts both.
"""
import datetime

import numpy as np

from . import h5lite

_MONTHS = ("JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV", "DEC")


def time_string(now=None):
    """'DD-MON-YYYY HH:MM:SS' as time_string builds it (phdf5.f90:1562-1618)."""
    t = now or datetime.datetime.now()
    return f"{t.day:02d}-{_MONTHS[t.month - 1]}-{t.year:4d} {t.hour:02d}:{t.minute:02d}:{t.second:02d}"


def flow_filename(time_step, final=False):
    """write_flow's file name (phdf5.f90:15-35): end.h5 for the final save,
    out.NNNNNN.h5 otherwise."""
    return "end.h5" if final else f"out.{int(time_step):06d}.h5"


def scalar_name(n):
    """Dataset name of scalar n (0-based): "TH"//char(ith+45) with ith = n+4."""
    return "TH" + chr(n + 4 + 45)


# --------------------------------------------------- y interpolations --
def g2gf(v, nx, nz):
    """g2gf (phdf5.f90:1543-1559) on a physical array [y][z][x] of Ny+2
    planes: V on the GY points -> V on the GYF points.  Returns a copy."""
    v = np.array(v, dtype=np.float64, copy=True)
    ny = v.shape[-3] - 2
    w = v[..., :, :nz, :nx]
    w[..., 1, :, :] = w[..., 2, :, :]
    w[..., ny + 1, :, :] = w[..., ny, :, :]
    w[..., 1:ny + 1, :, :] = 0.5 * (w[..., 1:ny + 1, :, :] + w[..., 2:ny + 2, :, :])
    return v


def gf2g(v, nx, nz):
    """gf2g (phdf5.f90:1500-1540), the inverse of g2gf's interior relation,
    with the zero-mass-flux wall values var(1) = -var(2), var(Ny+1) = -var(Ny).
    Returns a copy."""
    v = np.array(v, dtype=np.float64, copy=True)
    ny = v.shape[-3] - 2
    w = v[..., :, :nz, :nx]
    w[..., ny + 1, :, :] = w[..., ny, :, :]
    for y in range(ny, 0, -1):
        w[..., y, :, :] = 2 * w[..., y, :, :] - w[..., y + 1, :, :]
    w[..., 1, :, :] = -w[..., 2, :, :]
    w[..., ny + 1, :, :] = -w[..., ny, :, :]
    return v


def swapzy(v, nx, nz):
    """Physical [y][z][x] (Ny+2 planes) -> file array (Nz, Ny, Nx) of planes 1..Ny."""
    ny = v.shape[0] - 2
    return np.ascontiguousarray(v[1:ny + 1, :nz, :nx].transpose(1, 0, 2))


def swapyz(a, shape):
    """File array (Nz, Ny, Nx) -> physical [y][z][x] of the given shape, zero
    outside x < Nx, z < Nz, y = 1..Ny (swapyz clears the whole array first)."""
    out = np.zeros(shape)
    nz, ny, nx = a.shape
    out[1:ny + 1, :nz, :nx] = a.transpose(1, 0, 2)
    return out


# --------------------------------------------------------- write / read --
def write_flow(solver, path, time, time_step, save_flow_time=0.0, save_stats_time=0.0,
               save_movie_time=0.0, save_pressure=False, date=None, restore=True):
    """WriteHDF5 (phdf5.f90:913-1191) of the solver's state.

    With restore=True the solver's state afterwards is what the reference
    leaves behind: cu1..3, cth (and cp when saved) replaced by
    fft_xz_to_fourier of their physical copies, u2 after g2gf then gf2g
    (which sets its planes 1 and Ny+1 to -u2(2) and -u2(Ny))."""
    p = solver.p
    nx, ny, nz, nth = p.nx, p.ny, p.nz, p.n_th
    u1 = solver.get_physical("cu1")
    u2 = solver.get_physical("cu2")
    u3 = solver.get_physical("cu3")
    th = solver.get_physical("cth") if nth else None
    u2f = g2gf(u2, nx, nz)
    ts = {
        "attrs": {
            "Time": float(time),
            "Save_Flow_Time": float(save_flow_time),
            "Save_Stats_Time": float(save_stats_time),
            "Save_Movie_Time": float(save_movie_time),
            "Time_Step": np.int32(time_step),
        },
        "U": swapzy(u1, nx, nz),
        "V": swapzy(u2f, nx, nz),
        "W": swapzy(u3, nx, nz),
    }
    for n in range(nth):
        ts[scalar_name(n)] = swapzy(th[n], nx, nz)
    pr = None
    if save_pressure:
        pr = solver.get_physical("cp")
        ts["P"] = swapzy(pr, nx, nz)
    date = date if date is not None else time_string()
    if len(date) != 20:
        raise ValueError("date must be 20 characters ('DD-MON-YYYY HH:MM:SS')")
    tree = {
        "attrs": {
            "Resolution": np.array([nx, ny, nz], dtype=np.int32),
            "Date": np.frombuffer(date.encode("ascii"), dtype="S1"),
        },
        "Timestep": ts,
    }
    h5lite.write(path, tree)
    if restore:
        solver.set_physical("cu1", u1)
        solver.set_physical("cu2", gf2g(u2f, nx, nz))
        solver.set_physical("cu3", u3)
        if nth:
            solver.set_physical("cth", th)
        if pr is not None:
            solver.set_physical("cp", pr)


def read_flow_file(path):
    """The raw contents of a flow file: (attrs, timestep attrs, datasets)."""
    t = h5lite.read(path)
    if "Timestep" not in t:
        raise ValueError(f"{path}: no /Timestep group")
    ts = dict(t["Timestep"])
    tattrs = ts.pop("attrs", {})
    return t.get("attrs", {}), tattrs, ts


def read_flow(solver, path, create_new_th=None):
    """ReadHDF5 (phdf5.f90:1195-1459) into the solver's state.

    Reads U, V, W and each scalar n with create_new_th[n] false (the others
    keep their current values), and P when the file holds it.  Raises
    ValueError when the file's Resolution differs from the solver's
    (phdf5.f90:1283-1296).  Returns a dict with time, time_step,
    save_flow_time, save_stats_time, save_movie_time and read_pressure
    (compute_pressure = not read_pressure, ics.f90:446-448)."""
    p = solver.p
    nx, ny, nz, nth = p.nx, p.ny, p.nz, p.n_th
    attrs, tattrs, ds = read_flow_file(path)
    res = [int(v) for v in np.asarray(attrs.get("Resolution", [-1, -1, -1])).ravel()]
    if res != [nx, ny, nz]:
        raise ValueError(f"Error. File and program have different resolution. Program: {nx} {ny} {nz} File: {res}")
    new = [bool(x) for x in create_new_th] if create_new_th is not None else [False] * nth
    if len(new) != nth:
        raise ValueError(f"create_new_th must have {nth} entries")
    needed = ["U", "V", "W"] + [scalar_name(n) for n in range(nth) if not new[n]]
    for name in needed + (["P"] if "P" in ds else []):
        if name not in ds:
            raise ValueError(f"{path}: dataset /Timestep/{name} is missing")
        if tuple(ds[name].shape) != (nz, ny, nx):
            raise ValueError(f"{path}: /Timestep/{name} has shape {ds[name].shape}, expected {(nz, ny, nx)}")
    shape = (ny + 2, nz + 2, nx + 2)
    solver.set_physical("cu1", swapyz(ds["U"], shape))
    solver.set_physical("cu2", gf2g(swapyz(ds["V"], shape), nx, nz))
    solver.set_physical("cu3", swapyz(ds["W"], shape))
    if nth and not all(new):
        keep = solver.get_field("cth") if any(new) else None
        th = np.zeros((nth,) + shape)
        for n in range(nth):
            if not new[n]:
                th[n] = swapyz(ds[scalar_name(n)], shape)
        solver.set_physical("cth", th)
        if keep is not None:  # scalars created anew keep their current state
            cth = solver.get_field("cth")
            for n in range(nth):
                if new[n]:
                    cth[n] = keep[n]
            solver.set_field("cth", cth)
    read_pressure = "P" in ds
    if read_pressure:
        solver.set_physical("cp", swapyz(ds["P"], shape))
    return {
        "time": float(tattrs.get("Time", 0.0)),
        "time_step": int(tattrs.get("Time_Step", 0)),
        "save_flow_time": float(tattrs.get("Save_Flow_Time", 0.0)),
        "save_stats_time": float(tattrs.get("Save_Stats_Time", 0.0)),
        "save_movie_time": float(tattrs.get("Save_Movie_Time", 0.0)),
        "read_pressure": read_pressure,
    }


def write_start_file(path, U, V, W, TH=(), time=0.0, time_step=0, save_flow_time=0.0,
                     save_stats_time=0.0, save_movie_time=0.0):
    """A start.h5 built from host arrays, as py/setup/start.py:45-64 does:
    U, W and each TH of C shape (Nz, Ny, Nx) on the GYF points, V of shape
    (Nz, Ny+1, Nx) on the GY points, stored averaged to the GYF points
    ((V[:, :-1] + V[:, 1:]) / 2, start.py:17-18)."""
    U = np.asarray(U, dtype=np.float64)
    nz, ny, nx = U.shape
    V = np.asarray(V, dtype=np.float64)
    if V.shape != (nz, ny + 1, nx):
        raise ValueError(f"V must have shape {(nz, ny + 1, nx)}")
    ts = {
        "attrs": {
            "Time": float(time),
            "Save_Flow_Time": float(save_flow_time),
            "Save_Stats_Time": float(save_stats_time),
            "Save_Movie_Time": float(save_movie_time),
            "Time_Step": np.int32(time_step),
        },
        "U": U,
        "V": (V[:, :-1, :] + V[:, 1:, :]) / 2,
        "W": np.asarray(W, dtype=np.float64),
    }
    for n, t in enumerate(TH):
        ts[scalar_name(n)] = np.asarray(t, dtype=np.float64)
    for k in ("W",) + tuple(scalar_name(n) for n in range(len(TH))):
        if ts[k].shape != (nz, ny, nx):
            raise ValueError(f"{k} must have shape {(nz, ny, nx)}")
    h5lite.write(path, {"attrs": {"Resolution": np.array([nx, ny, nz], dtype=np.int32)}, "Timestep": ts})

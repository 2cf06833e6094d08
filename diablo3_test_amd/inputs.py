"""DIABLO's run inputs: grid_def.all, grid.h5 and input.toml.

* grid_def.all holds the compile-time sizes Nx, Ny, Nz, N_th
  (parameters.f90:20, py/setup/setup.py:16-47);
* grid.h5 holds /grids/y, the Ny+1 base-grid points GY(1..Ny+1)
  (domain.f90:429-483), made by py/setup/create_grid.py:22-72 (restated in
  grid.create_grid) and written by write_grid_h5;
* input.toml holds the run parameters (parameters.f90:299-405, schema of
  test_case/input.toml), turned into the module globals by read_inputs
  (parameters.f90:91-105) and set_parameters (parameters.f90:241-281).

`channel_params` turns these into the parameter block of the channel solver.
"""
import math
import re

import numpy as np

from . import h5lite

INPUT_VERSION = "3.5"  # parameters.f90:10; the run stops on any other VERSION


# ------------------------------------------------------------- grid_def --
def read_grid_def(path):
    """(Nx, Ny, Nz, N_th) from a grid_def.all file: lines `name = value`,
    names matched case-insensitively, `#` and `!` start comments."""
    found = {}
    pat = re.compile(r"\b(N_?TH|NX|NY|NZ)\s*=\s*([0-9]+)", re.IGNORECASE)
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].split("!", 1)[0]
            for m in pat.finditer(line):
                found[m.group(1).upper().replace("_", "")] = int(m.group(2))
    missing = [k for k in ("NX", "NY", "NZ", "NTH") if k not in found]
    if missing:
        raise ValueError(f"{path}: no value for {', '.join(missing)}")
    return found["NX"], found["NY"], found["NZ"], found["NTH"]


# -------------------------------------------------------------- grid.h5 --
def read_grid_h5(path):
    """GY(1..Ny+1) from grid.h5:/grids/y (domain.f90:429-483)."""
    y = np.asarray(h5lite.read_dataset(path, "grids/y"), dtype=np.float64)
    if y.ndim != 1:
        raise ValueError(f"{path}: /grids/y must be one-dimensional")
    return y


def write_grid_h5(path, gy):
    """Write GY(1..Ny+1) as grid.h5:/grids/y (create_grid.py:159-161)."""
    h5lite.write(path, {"grids": {"y": np.asarray(gy, dtype=np.float64)}})


# ----------------------------------------------------------- input.toml --
def _load_toml(path):
    try:
        import tomllib as toml
    except ImportError:  # Python < 3.11
        import tomli as toml
    with open(path, "rb") as f:
        return toml.load(f)


def _get(table, section, key, kind):
    """Typed lookup with the checks of read_input_toml's get_*_from_table
    wrappers (parameters.f90:407-480): a missing key or a wrong type stops
    the run.  An integer is accepted where a float is read."""
    where = f"{section}.{key}" if section else key
    if key not in table:
        raise ValueError(f"input.toml: {where} is missing")
    v = table[key]
    if kind == "float":
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise ValueError(f"input.toml: {where} must be a float")
        return float(v)
    if kind == "int":
        if isinstance(v, bool) or not isinstance(v, int):
            raise ValueError(f"input.toml: {where} must be an integer")
        return v
    if kind == "bool":
        if not isinstance(v, bool):
            raise ValueError(f"input.toml: {where} must be a boolean")
        return v
    if kind == "str":
        if not isinstance(v, str):
            raise ValueError(f"input.toml: {where} must be a string")
        return v
    if kind == "float3":
        if not isinstance(v, list) or len(v) != 3:
            raise ValueError(f"input.toml: {where} must be an array of 3 floats")
        return [_get({"x": e}, section, "x", "float") for e in v]
    raise AssertionError(kind)


def _section(root, name):
    t = root.get(name)
    if not isinstance(t, dict):
        raise ValueError(f"input.toml: table [{name}] is missing")
    return t


def read_input_toml(path, n_th):
    """The module globals that read_input_toml (parameters.f90:299-405) and
    read_inputs (parameters.f90:91-105) set, as a dict with the Fortran
    names in lower case.  n_th is the compile-time scalar count (grid_def);
    the [[SCALARS]] array must have exactly that many entries.

    The TOML keys V_BC_* set the spanwise (w) and W_BC_* the wall-normal (v)
    velocity conditions (parameters.f90:373-382): DIABLO's TOML names the
    vertical direction W while the code calls it v.
    """
    root = _load_toml(path)
    c = {}
    c["version"] = _get(root, "", "VERSION", "str")
    if c["version"] != INPUT_VERSION:
        raise ValueError("Wrong input version")
    t = _section(root, "SCHEME")
    c["flavor"] = _get(t, "SCHEME", "FLAVOUR", "str")
    c["use_les"] = _get(t, "SCHEME", "USE_LES", "bool")
    c["les_model_type"] = _get(t, "SCHEME", "LES_MODEL_TYPE", "int")
    c["beta"] = _get(t, "SCHEME", "BETA", "float")
    c["time_ad_meth"] = 1
    t = _section(root, "PHYSICAL")
    for k in ("LX", "LY", "LZ", "RE", "NU_V_SCALE", "RO", "DELTA"):
        c[k.lower()] = _get(t, "PHYSICAL", k, "float")
    c["grav_x"], c["grav_y"], c["grav_z"] = _get(t, "PHYSICAL", "GRAV", "float3")
    t = _section(root, "TIMESTEPPING")
    for k in ("WALL_TIME_LIMIT", "TIME_LIMIT", "DELTA_T", "CFL"):
        c[k.lower()] = _get(t, "TIMESTEPPING", k, "float")
    c["variable_dt"] = _get(t, "TIMESTEPPING", "VARIABLE_DT", "bool")
    c["update_dt"] = _get(t, "TIMESTEPPING", "UPDATE_DT", "int")
    t = _section(root, "OUTPUT")
    c["verbosity"] = _get(t, "OUTPUT", "VERBOSITY", "int")
    for k in ("SAVE_FLOW_DT", "SAVE_STATS_DT", "SAVE_MOVIE_DT"):
        c[k.lower()] = _get(t, "OUTPUT", k, "float")
    c["xcmovie"], c["ycmovie"], c["zcmovie"] = _get(t, "OUTPUT", "MOVIE", "float3")
    t = _section(root, "INITIAL_CONDITIONS")
    c["create_new_flow"] = _get(t, "INITIAL_CONDITIONS", "CREATE_NEW_FLOW", "bool")
    c["reset_time"] = _get(t, "INITIAL_CONDITIONS", "RESET_TIME", "bool")
    c["ic_type"] = _get(t, "INITIAL_CONDITIONS", "IC_TYPE", "int")
    c["kick"] = _get(t, "INITIAL_CONDITIONS", "KICK", "float")
    c["physical_noise"] = _get(t, "INITIAL_CONDITIONS", "PHYSICAL_NOISE", "bool")
    t = _section(root, "FORCING")
    c["f_type"] = _get(t, "FORCING", "F_TYPE", "int")
    for k in ("UBULK0", "PX0", "OMEGA0", "AMP_OMEGA0", "FORCE_START"):
        c[k.lower()] = _get(t, "FORCING", k, "float")
    t = _section(root, "VELOCITY_BCS")
    # (TOML key prefix, code variable) -- the V/W swap of parameters.f90:373-382
    for toml_name, var in (("U", "u"), ("V", "w"), ("W", "v")):
        for side in ("YMIN", "YMAX"):
            c[f"{var}_bc_{side.lower()}"] = _get(t, "VELOCITY_BCS", f"{toml_name}_BC_{side}", "int")
            c[f"{var}_bc_{side.lower()}_c1"] = _get(t, "VELOCITY_BCS", f"{toml_name}_BC_{side}_C1", "float")
    scal = root.get("SCALARS", [])
    if not isinstance(scal, list):
        raise ValueError("input.toml: SCALARS must be an array of tables ([[SCALARS]])")
    if len(scal) != n_th:
        raise ValueError(f"Error: {n_th} scalars defined but {len(scal)} scalars found in input.toml")
    keys = [("CREATE_FLOW_TH", "create_new_th", "bool"), ("FILTER_TH", "filter_th", "bool"),
            ("FILTER_INT", "filter_int", "int"), ("RI", "ri", "float"), ("PR", "pr", "float"),
            ("TH_BC_YMIN", "th_bc_ymin", "int"), ("TH_BC_YMIN_C1", "th_bc_ymin_c1", "float"),
            ("TH_BC_YMAX", "th_bc_ymax", "int"), ("TH_BC_YMAX_C1", "th_bc_ymax_c1", "float")]
    for _, var, _ in keys:
        c[var] = []
    for n, s in enumerate(scal):
        for key, var, kind in keys:
            c[var].append(_get(s, f"SCALARS[{n + 1}]", key, kind))
    # read_inputs (parameters.f90:91-105)
    c["nu"] = 1.0 / c["re"]
    c["ro_inv"] = 1.0 / c["ro"]
    c["max_dt"] = c["delta_t"]
    return c


def set_parameters(c, n_th):
    """set_parameters (parameters.f90:241-281) on a dict from read_input_toml:
    the gravity-current compensation of the spanwise wall stress and the
    background scalar gradients dTHdX, dTHdZ of the front initial conditions,
    the transient top-wall value and the averaging flag.  Returns a new dict."""
    c = dict(c)
    ic = c["ic_type"]
    dthdx = [0.0] * n_th
    dthdz = [0.0] * n_th
    if ic in (4, 5):  # infinite front
        if c["w_bc_ymin"] == 1:
            c["w_bc_ymin_c1"] -= 1.0
        if c["w_bc_ymax"] == 1:
            c["w_bc_ymax_c1"] -= 1.0
        if n_th:
            dthdx[0] = c["ro_inv"] / c["delta"]
    elif ic in (6, 7, 8):  # finite front
        if c["w_bc_ymin"] == 1:
            c["w_bc_ymin_c1"] -= 2.0 * c["delta"] / c["lx"]
        if c["w_bc_ymax"] == 1:
            c["w_bc_ymax_c1"] -= 2.0 * c["delta"] / c["lx"]
        if n_th:
            dthdx[0] = 2.0 / c["lx"] * c["ro_inv"]
    c["dthdx"] = dthdx
    c["dthdz"] = dthdz
    c["w_bc_ymax_c1_transient"] = c["w_bc_ymax_c1"]
    c["homogeneousx"] = ic not in (5, 6, 7, 8)
    return c


def channel_params(c, nx, ny, nz, n_th):
    """The solver's parameter dict from set_parameters' output and the
    grid_def sizes."""
    if not math.isfinite(c["delta_t"]) or c["delta_t"] <= 0:
        raise ValueError("DELTA_T must be positive")
    return dict(
        nx=nx, ny=ny, nz=nz, n_th=n_th,
        lx=c["lx"], lz=c["lz"], re=c["re"], nu_v_scale=c["nu_v_scale"], beta=c["beta"],
        ro_inv=c["ro_inv"], delta=c["delta"],
        grav_x=c["grav_x"], grav_y=c["grav_y"], grav_z=c["grav_z"],
        delta_t=c["delta_t"], f_type=c["f_type"], les_model_type=c["les_model_type"],
        px0=c["px0"], omega0=c["omega0"], amp_omega0=c["amp_omega0"], force_start=c["force_start"],
        u_bc_ymin=c["u_bc_ymin"], v_bc_ymin=c["v_bc_ymin"], w_bc_ymin=c["w_bc_ymin"],
        u_bc_ymax=c["u_bc_ymax"], v_bc_ymax=c["v_bc_ymax"], w_bc_ymax=c["w_bc_ymax"],
        u_bc_ymin_c1=c["u_bc_ymin_c1"], v_bc_ymin_c1=c["v_bc_ymin_c1"], w_bc_ymin_c1=c["w_bc_ymin_c1"],
        u_bc_ymax_c1=c["u_bc_ymax_c1"], v_bc_ymax_c1=c["v_bc_ymax_c1"], w_bc_ymax_c1=c["w_bc_ymax_c1"],
        pr=list(c["pr"]), th_bc_ymin=list(c["th_bc_ymin"]), th_bc_ymax=list(c["th_bc_ymax"]),
        th_bc_ymin_c1=list(c["th_bc_ymin_c1"]), th_bc_ymax_c1=list(c["th_bc_ymax_c1"]),
        dthdx=list(c["dthdx"]), dthdz=list(c["dthdz"]),
        use_les=int(c["use_les"]), create_new_flow=int(c["create_new_flow"]),
        # flavor == 'Front' is a blank-padded Fortran comparison (forcing.f90:50)
        flavor_front=int(c["flavor"].rstrip() == "Front"), ly=c["ly"],
    )


def read_run(directory):
    """grid_def.all, grid.h5 and input.toml of a run directory ->
    (solver parameter dict, gy, full input dict after set_parameters)."""
    import os

    nx, ny, nz, n_th = read_grid_def(os.path.join(directory, "grid_def.all"))
    gy = read_grid_h5(os.path.join(directory, "grid.h5"))
    if gy.size != ny + 1:
        raise ValueError(f"grid.h5 holds {gy.size} points, expected Ny+1 = {ny + 1}")
    c = set_parameters(read_input_toml(os.path.join(directory, "input.toml"), n_th), n_th)
    return channel_params(c, nx, ny, nz, n_th), gy, c

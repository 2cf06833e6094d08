"""ctypes binding of the C-ABI in include/diablo_gpu.h.

The reference's operator API for this path is a set of argument-less Fortran
module procedures working on module globals (for/channel.f90:10 rk_chan_1,
:946 rem_div_chan; for/boundary.f90:824 apply_BC_vel_mpi_post; the driver loop
for/diablo.f90:98-111).  ``ChannelSolver`` keeps those names; the globals they
read become the ``Params`` structure plus the y-grid, and the state arrays of
for/flow.f90:20-25, 50-51 are exchanged in the reference's own memory layout:

* spectral fields (cu1, cu2, cu3, cp, cth, cf1..3, cfth): complex128 arrays of
  shape (Ny+2, Nz+2, Nx/2+1) -- the Fortran array cu1(0:Nx/2, 0:Nz+1, 0:Ny+1)
  seen from C;
* physical fields (nu_t, kappa_t): float64 arrays of shape (Ny+2, Nz+2, Nx+2);
* per-scalar fields (cth, cfth, kappa_t) carry a leading N_th axis.

The HIP library must be built (``diablo3_test_amd.build.build_library``);
loading fails loudly when it is missing.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(HERE, "libdgpu.so")

MAX_TH = 16

DGPU_OK = 0
ERRORS = {-1: "EINVAL", -2: "EHIP", -3: "ENOMEM", -4: "EUNSUP", -5: "ESTATE"}

PRECISIONS = {"fp64": 0, "mixed": 1}

FIELD_IDS = {
    "cu1": 0,
    "cu2": 1,
    "cu3": 2,
    "cp": 3,
    "cth": 4,
    "cf1": 5,
    "cf2": 6,
    "cf3": 7,
    "cfth": 8,
    "nu_t": 9,
    "kappa_t": 10,
}
PER_SCALAR = ("cth", "cfth", "kappa_t")
PHYSICAL = ("nu_t", "kappa_t")


class Params(ctypes.Structure):
    """Mirror of ``dgpu_params`` (include/diablo_gpu.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32),
        ("ny", ctypes.c_int32),
        ("nz", ctypes.c_int32),
        ("n_th", ctypes.c_int32),
        ("lx", ctypes.c_double),
        ("lz", ctypes.c_double),
        ("re", ctypes.c_double),
        ("nu_v_scale", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("ro_inv", ctypes.c_double),
        ("delta", ctypes.c_double),
        ("grav_x", ctypes.c_double),
        ("grav_y", ctypes.c_double),
        ("grav_z", ctypes.c_double),
        ("delta_t", ctypes.c_double),
        ("f_type", ctypes.c_int32),
        ("les_model_type", ctypes.c_int32),
        ("px0", ctypes.c_double),
        ("omega0", ctypes.c_double),
        ("amp_omega0", ctypes.c_double),
        ("force_start", ctypes.c_double),
        ("u_bc_ymin", ctypes.c_int32),
        ("v_bc_ymin", ctypes.c_int32),
        ("w_bc_ymin", ctypes.c_int32),
        ("u_bc_ymax", ctypes.c_int32),
        ("v_bc_ymax", ctypes.c_int32),
        ("w_bc_ymax", ctypes.c_int32),
        ("u_bc_ymin_c1", ctypes.c_double),
        ("v_bc_ymin_c1", ctypes.c_double),
        ("w_bc_ymin_c1", ctypes.c_double),
        ("u_bc_ymax_c1", ctypes.c_double),
        ("v_bc_ymax_c1", ctypes.c_double),
        ("w_bc_ymax_c1", ctypes.c_double),
        ("pr", ctypes.c_double * MAX_TH),
        ("th_bc_ymin", ctypes.c_int32 * MAX_TH),
        ("th_bc_ymax", ctypes.c_int32 * MAX_TH),
        ("th_bc_ymin_c1", ctypes.c_double * MAX_TH),
        ("th_bc_ymax_c1", ctypes.c_double * MAX_TH),
        ("dthdx", ctypes.c_double * MAX_TH),
        ("dthdz", ctypes.c_double * MAX_TH),
        ("use_les", ctypes.c_int32),
        ("create_new_flow", ctypes.c_int32),
        ("flavor_front", ctypes.c_int32),
        ("ly", ctypes.c_double),
    ]

    @classmethod
    def from_dict(cls, d):
        p = cls()
        for k, v in d.items():
            f = getattr(p, k)
            if isinstance(f, ctypes.Array):
                arr = list(v) + [0] * (len(f) - len(v))
                for i, x in enumerate(arr):
                    f[i] = x
            else:
                setattr(p, k, v)
        return p

    def to_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, ctypes.Array) else v
        return out


class CourantParams(ctypes.Structure):
    """Mirror of ``dgpu_courant_params`` (include/diablo_gpu.h)."""

    _fields_ = [
        ("cfl", ctypes.c_double),
        ("max_dt", ctypes.c_double),
        ("ly", ctypes.c_double),
        ("time_limit", ctypes.c_double),
        ("save_stats_time", ctypes.c_double),
        ("save_flow_time", ctypes.c_double),
        ("update_dt", ctypes.c_int32),
        ("front", ctypes.c_int32),
    ]


class DgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_LIB = {}


def _declare(lib, prefix, with_precision=True):
    P = ctypes.POINTER
    h = ctypes.c_void_p
    create = getattr(lib, prefix + "create")
    if with_precision:
        create.argtypes = [P(Params), P(ctypes.c_double), ctypes.c_int, P(h)]
    else:
        create.argtypes = [P(Params), P(ctypes.c_double), P(h)]
    create.restype = ctypes.c_int
    for name, args in [
        ("destroy", [h]),
        ("set_field", [h, ctypes.c_int, P(ctypes.c_double)]),
        ("get_field", [h, ctypes.c_int, P(ctypes.c_double)]),
        ("get_physical", [h, ctypes.c_int, P(ctypes.c_double)]),
        ("set_physical", [h, ctypes.c_int, P(ctypes.c_double)]),
        ("poisson_p", [h, P(ctypes.c_double)]),
        ("set_courant", [h, P(CourantParams)]),
        ("courant", [h, ctypes.c_double]),
        ("get_delta_t", [h, P(ctypes.c_double)]),
        ("set_delta_t", [h, ctypes.c_double]),
        ("apply_bc_vel_post", [h, ctypes.c_double]),
        ("rem_div", [h]),
        ("rk_substep", [h, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
        ("step", [h, ctypes.c_int, P(ctypes.c_double), P(ctypes.c_int32)]),
        ("synchronize", [h]),
    ]:
        fn = getattr(lib, prefix + name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    le = getattr(lib, prefix + "last_error")
    le.argtypes = [h]
    le.restype = ctypes.c_char_p
    fd = getattr(lib, prefix + "field_doubles")
    fd.argtypes = [h, ctypes.c_int]
    fd.restype = ctypes.c_int64
    return lib


def load_library(path=None):
    """Load the HIP library (``libdgpu.so``).  Raises if it has not been built."""
    path = path or os.environ.get("DGPU_LIB", DEFAULT_LIB)
    if path not in _LIB:
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} not found: build it with diablo3_test_amd.build.build_library()"
            )
        lib = _declare(ctypes.CDLL(path), "dgpu_")
        lib.dgpu_timing_enable.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.dgpu_timing_enable.restype = ctypes.c_int
        P = ctypes.POINTER
        lib.dgpu_timing_read.argtypes = [ctypes.c_void_p, P(ctypes.c_double), P(ctypes.c_double),
                                         P(ctypes.c_int64)]
        lib.dgpu_timing_read.restype = ctypes.c_int
        lib.dgpu_timing_read_intermediate.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        lib.dgpu_timing_read_intermediate.restype = ctypes.c_int
        lib.dgpu_kernel_class_name.argtypes = [ctypes.c_int]
        lib.dgpu_kernel_class_name.restype = ctypes.c_char_p
        h = ctypes.c_void_p
        d = P(ctypes.c_double)
        for name, args in [
            ("adj_set_field", [h, ctypes.c_int, d]),
            ("adj_get_field", [h, ctypes.c_int, d]),
            ("adj_zero", [h]),
            ("adj_rk_substep", [h, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            ("adj_add_mean", [h, ctypes.c_int, d]),
            ("mean_profile", [h, ctypes.c_int, d]),
            ("checkpoint", [h, ctypes.c_int, ctypes.c_int]),
            ("check_state", [h]),
            ("set_field_device", [h, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
            ("get_field_device", [h, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
            ("adj_set_field_device", [h, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
            ("adj_get_field_device", [h, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
            ("check_state_noalloc", [h]),
        ]:
            fn = getattr(lib, "dgpu_" + name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _LIB[path] = lib
    return _LIB[path]


NCLASS = 13


def spectral_shape(p):
    return (p.ny + 2, p.nz + 2, p.nx // 2 + 1)


def physical_shape(p):
    return (p.ny + 2, p.nz + 2, p.nx + 2)


class ChannelSolver:
    """One channel-flow state on one GPU, advanced by rk_chan_1 substeps.

    ``params`` is a ``Params`` or a dict of its fields; ``gy`` holds the Ny+1
    values of grid.h5:/grids/y.  ``lib``/``prefix`` select another library
    exporting the same C-ABI under another symbol prefix.
    """

    def __init__(self, params, gy, precision="fp64", lib=None, prefix="dgpu_"):
        self.p = params if isinstance(params, Params) else Params.from_dict(params)
        self.gy = np.ascontiguousarray(gy, dtype=np.float64)
        if self.gy.size != self.p.ny + 1:
            raise ValueError(f"gy must hold Ny+1 = {self.p.ny + 1} values, got {self.gy.size}")
        self.lib = lib if lib is not None else load_library()
        self.prefix = prefix
        self._h = ctypes.c_void_p()
        gyp = self.gy.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        if prefix == "dgpu_":
            rc = self._fn("create")(ctypes.byref(self.p), gyp, PRECISIONS[precision], ctypes.byref(self._h))
        else:
            rc = self._fn("create")(ctypes.byref(self.p), gyp, ctypes.byref(self._h))
        if rc != DGPU_OK:
            raise DgpuError(rc, f"{prefix}create failed (bad parameters or device error)")
        self.time = 0.0
        self.time_step = 0
        self._courant = None

    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc, what):
        if rc != DGPU_OK:
            msg = self._fn("last_error")(self._h)
            raise DgpuError(rc, f"{what}: {msg.decode() if msg else ''}")

    # ---- state exchange in the reference layout
    def _shape(self, name):
        base = physical_shape(self.p) if name in PHYSICAL else spectral_shape(self.p)
        if name in PER_SCALAR:
            base = (self.p.n_th,) + base
        return base

    def set_field(self, name, arr):
        fid = FIELD_IDS[name]
        shape = self._shape(name)
        dt = np.float64 if name in PHYSICAL else np.complex128
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt).reshape(shape))
        self._check(
            self._fn("set_field")(self._h, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            f"set_field({name})",
        )

    def get_field(self, name):
        fid = FIELD_IDS[name]
        shape = self._shape(name)
        dt = np.float64 if name in PHYSICAL else np.complex128
        a = np.zeros(shape, dtype=dt)
        self._check(
            self._fn("get_field")(self._h, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            f"get_field({name})",
        )
        return a

    # ---- the same exchange with torch tensors in device memory
    def _device_tensor(self, name, t):
        import torch

        dt = torch.float64 if name in PHYSICAL else torch.complex128
        if not (isinstance(t, torch.Tensor) and t.is_cuda):
            raise TypeError(f"{name}: a device tensor is required")
        if t.dtype != dt or tuple(t.shape) != self._shape(name) or not t.is_contiguous():
            raise ValueError(f"{name}: expected a contiguous {dt} tensor of shape {self._shape(name)}, "
                             f"got {t.dtype} {tuple(t.shape)}")
        return t

    @staticmethod
    def _stream(t, stream):
        import torch

        s = stream if stream is not None else torch.cuda.current_stream(t.device)
        return ctypes.c_void_p(s.cuda_stream)

    def _new_device(self, name, device):
        import torch

        dt = torch.float64 if name in PHYSICAL else torch.complex128
        return torch.empty(self._shape(name), dtype=dt, device=device)

    def set_field_device(self, name, t, stream=None):
        """set_field from a device tensor (its data_ptr), ordered on `stream`
        (default: torch's current stream on the tensor's device)."""
        t = self._device_tensor(name, t)
        self._check(self.lib.dgpu_set_field_device(self._h, FIELD_IDS[name], ctypes.c_void_p(t.data_ptr()),
                                                   self._stream(t, stream)), f"set_field_device({name})")

    def get_field_device(self, name, out=None, stream=None, device="cuda"):
        """get_field into a device tensor (new, or `out`), ordered on `stream`."""
        t = self._device_tensor(name, out if out is not None else self._new_device(name, device))
        self._check(self.lib.dgpu_get_field_device(self._h, FIELD_IDS[name], ctypes.c_void_p(t.data_ptr()),
                                                   self._stream(t, stream)), f"get_field_device({name})")
        return t

    def adj_set_field_device(self, name, t, stream=None):
        t = self._device_tensor(name, t)
        self._check(self.lib.dgpu_adj_set_field_device(self._h, FIELD_IDS[name], ctypes.c_void_p(t.data_ptr()),
                                                       self._stream(t, stream)), f"adj_set_field_device({name})")

    def adj_get_field_device(self, name, out=None, stream=None, device="cuda"):
        t = self._device_tensor(name, out if out is not None else self._new_device(name, device))
        self._check(self.lib.dgpu_adj_get_field_device(self._h, FIELD_IDS[name], ctypes.c_void_p(t.data_ptr()),
                                                       self._stream(t, stream)), f"adj_get_field_device({name})")
        return t

    def set_state(self, state):
        for k, v in state.items():
            self.set_field(k, v)

    def get_state(self, names=("cu1", "cu2", "cu3", "cp", "cth")):
        return {k: self.get_field(k) for k in names if not (k in PER_SCALAR and self.p.n_th == 0)}

    def _phys_shape(self, name):
        base = physical_shape(self.p)
        return ((self.p.n_th,) + base) if name in PER_SCALAR else base

    def get_physical(self, name):
        """fft_xz_to_physical of spectral field `name` (cu1..3, cp, cth), computed
        on the device; float64 (Ny+2, Nz+2, Nx+2) arrays ((n_th, ...) for cth)."""
        if name not in ("cu1", "cu2", "cu3", "cp", "cth"):
            raise ValueError(f"no physical copy of {name}")
        a = np.zeros(self._phys_shape(name))
        self._check(
            self._fn("get_physical")(self._h, FIELD_IDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            f"get_physical({name})",
        )
        return a

    def set_physical(self, name, arr):
        """Spectral field `name` = fft_xz_to_fourier of physical `arr` on all planes."""
        if name not in ("cu1", "cu2", "cu3", "cp", "cth"):
            raise ValueError(f"no physical copy of {name}")
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64).reshape(self._phys_shape(name)))
        self._check(
            self._fn("set_physical")(self._h, FIELD_IDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            f"set_physical({name})",
        )

    # ---- the reference's procedures
    def apply_bc_vel_mpi_post(self, time=None):
        """apply_BC_vel_mpi_post (for/boundary.f90:824-838)."""
        t = self.time if time is None else time
        self._check(self._fn("apply_bc_vel_post")(self._h, t), "apply_bc_vel_post")

    def rem_div_chan(self):
        """rem_div_chan (for/channel.f90:946-1023) without the pressure update."""
        self._check(self._fn("rem_div")(self._h), "rem_div")

    def pre_first_step(self, compute_pressure=False, ri=None):
        """pre_first_step (for/advance.f90:17-44): wall values, projection,
        poisson_p_chan when compute_pressure (start files without P; ri holds
        Ri(1..n_th)), and courant when variable_dt is on (set_courant)."""
        self.apply_bc_vel_mpi_post()
        self.rem_div_chan()
        if compute_pressure:
            self.poisson_p_chan(ri)
        if self._courant is not None:
            self.courant()

    def poisson_p_chan(self, ri=None):
        """poisson_p_chan (for/channel.f90:1026-1234): cp from the velocity and
        the scalars; ri = Ri(1..n_th) (parameters.f90:73)."""
        n = self.p.n_th
        r = np.zeros(max(n, 1))
        if n:
            if ri is None or len(ri) != n:
                raise ValueError(f"ri must hold {n} Richardson numbers")
            r[:n] = np.asarray(ri, dtype=np.float64)
        self._check(self._fn("poisson_p")(self._h, r.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    "poisson_p")

    def set_courant(self, cfl=None, max_dt=None, update_dt=1, ly=0.0, time_limit=1e300,
                    save_stats_time=1e300, save_flow_time=1e300, front=False, off=False):
        """Turn variable_dt on with courant's inputs (advance.f90:47-158), or
        off with off=True.  max_dt defaults to the current delta_t (as
        parameters.f90:103 sets it from DELTA_T); the output times steer the
        step towards the next save as courant does."""
        if off:
            self._courant = None
            self._check(self._fn("set_courant")(self._h, None), "set_courant")
            return
        c = CourantParams(cfl=float(cfl), max_dt=float(self.delta_t if max_dt is None else max_dt),
                          ly=float(ly), time_limit=float(time_limit), save_stats_time=float(save_stats_time),
                          save_flow_time=float(save_flow_time), update_dt=int(update_dt), front=int(bool(front)))
        self._check(self._fn("set_courant")(self._h, ctypes.byref(c)), "set_courant")
        self._courant = c

    def courant(self, time=None):
        """courant as pre_first_step calls it, on the current state."""
        t = self.time if time is None else time
        self._check(self._fn("courant")(self._h, float(t)), "courant")

    @property
    def delta_t(self):
        d = ctypes.c_double()
        self._check(self._fn("get_delta_t")(self._h, ctypes.byref(d)), "get_delta_t")
        return d.value

    @delta_t.setter
    def delta_t(self, dt):
        self._check(self._fn("set_delta_t")(self._h, float(dt)), "set_delta_t")

    def rk_chan_1(self, rk_step, time=None, time_step=None):
        """One RK3 substep (for/channel.f90:10-917)."""
        t = self.time if time is None else time
        ts = self.time_step if time_step is None else time_step
        self._check(self._fn("rk_substep")(self._h, int(rk_step), float(t), int(ts)), "rk_substep")

    def step(self, nsteps=1):
        """nsteps passes of the driver loop (for/diablo.f90:98-108)."""
        t = ctypes.c_double(self.time)
        ts = ctypes.c_int32(self.time_step)
        self._check(self._fn("step")(self._h, int(nsteps), ctypes.byref(t), ctypes.byref(ts)), "step")
        self.time = t.value
        self.time_step = ts.value

    def timing_enable(self, on=True):
        """Bracket every kernel launch with HIP events (resets the sums)."""
        self._check(self.lib.dgpu_timing_enable(self._h, 1 if on else 0), "timing_enable")

    def timing_read(self):
        """{kernel class: (ms, algorithmic bytes, launches)} since timing_enable."""
        ms = (ctypes.c_double * NCLASS)()
        by = (ctypes.c_double * NCLASS)()
        ct = (ctypes.c_int64 * NCLASS)()
        self._check(self.lib.dgpu_timing_read(self._h, ms, by, ct), "timing_read")
        out = {}
        for c in range(NCLASS):
            name = self.lib.dgpu_kernel_class_name(c).decode()
            out[name] = (ms[c], by[c], ct[c])
        return out

    def timing_read_intermediate(self):
        """{kernel class: x-spectra bytes} since timing_enable: the part of the
        algorithmic bytes a one-pass 2-D transform would not move."""
        tb = (ctypes.c_double * NCLASS)()
        self._check(self.lib.dgpu_timing_read_intermediate(self._h, tb), "timing_read_intermediate")
        return {self.lib.dgpu_kernel_class_name(c).decode(): tb[c] for c in range(NCLASS)}

    def synchronize(self):
        self._check(self._fn("synchronize")(self._h), "synchronize")

    def check_state(self):
        """True when the state (cu1..3, cp, cth, the RK carry and, with LES,
        nu_t/kappa_t) is finite; False when dgpu_check_state reports
        DGPU_ESTATE (``self.state_error`` then names the fields)."""
        return self._check_state(self.lib.dgpu_check_state)

    def check_state_noalloc(self):
        """check_state through dgpu_check_state_noalloc (no device allocation)."""
        return self._check_state(self.lib.dgpu_check_state_noalloc)

    def _check_state(self, fn):
        rc = fn(self._h)
        if rc == -5:
            self.state_error = self.lib.dgpu_last_error(self._h).decode()
            return False
        self._check(rc, "check_state")
        return True

    # ---- adjoint of the FP64 DNS substep (HIP library only)
    def adj_set_field(self, name, arr):
        fid = FIELD_IDS[name]
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.complex128).reshape(self._shape(name)))
        self._check(self.lib.dgpu_adj_set_field(self._h, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    f"adj_set_field({name})")

    def adj_get_field(self, name):
        fid = FIELD_IDS[name]
        a = np.zeros(self._shape(name), dtype=np.complex128)
        self._check(self.lib.dgpu_adj_get_field(self._h, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
                    f"adj_get_field({name})")
        return a

    def adj_set_state(self, state):
        for k, v in state.items():
            self.adj_set_field(k, v)

    def adj_get_state(self, names=("cu1", "cu2", "cu3", "cp", "cth", "cf1", "cf2", "cf3", "cfth")):
        return {k: self.adj_get_field(k) for k in names if not (k in PER_SCALAR and self.p.n_th == 0)}

    def adj_zero(self):
        self._check(self.lib.dgpu_adj_zero(self._h), "adj_zero")

    def adj_rk_chan_1(self, rk_step, time=None, time_step=None):
        """Adjoint of rk_chan_1: forward state = substep input, adjoint state =
        gradient w.r.t. the substep output -> gradient w.r.t. the input."""
        t = self.time if time is None else time
        ts = self.time_step if time_step is None else time_step
        self._check(self.lib.dgpu_adj_rk_substep(self._h, int(rk_step), float(t), int(ts)), "adj_rk_substep")

    def adj_add_mean(self, name, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.p.ny + 2
        self._check(self.lib.dgpu_adj_add_mean(self._h, FIELD_IDS[name],
                                               w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "adj_add_mean")

    def mean_profile(self, name="cu1"):
        out = np.zeros(self.p.ny + 2)
        self._check(self.lib.dgpu_mean_profile(self._h, FIELD_IDS[name],
                                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "mean_profile")
        return out

    def checkpoint(self, slot, save=True, light=False):
        """Save (or restore) the forward state in device slot `slot`; a light
        save keeps only the fields the adjoint substep reads (cu1..3, cth)."""
        mode = (2 if light else 1) if save else 0
        self._check(self.lib.dgpu_checkpoint(self._h, int(slot), mode), "checkpoint")

    def close(self):
        if self._h:
            self._fn("destroy")(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

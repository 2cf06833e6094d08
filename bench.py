"""Throughput of the channel RK3/CN step on one MI355X.

A step is one pass of the driver loop (for/diablo.f90:98-108): three
rk_chan_1 substeps over the whole field, which stays resident in HBM.
The metric is Mpts.RK3-steps/s = Nx*Ny*Nz * steps / seconds / 1e6.

Beside the GPU number the line carries
* roofline: the dominant kernel class's algorithmic HBM bytes over its
  kernel time, timed live with HIP events on the library's stream;
* cpu_baseline: the CPU oracle (C++ FP64 restatement, OpenMP) timed on the
  host cores on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts·RK3-steps/s at 512×257×512: FP64 plain vs mixed+adjoint; % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(p, gy, st, npts):
    """One RK3 step (three rk_chan_1 substeps) of the same state on the CPU oracle."""
    import ctypes

    from diablo3_test_amd.binding import ChannelSolver, _declare

    lib = _declare(ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so")), "orc_", with_precision=False)
    o = ChannelSolver(p, gy, lib=lib, prefix="orc_")
    o.set_state(st)
    o.pre_first_step()
    t0 = time.perf_counter()
    o.step(1)
    dt = time.perf_counter() - t0
    o.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": npts / dt / 1e6,
        "unit": "Mpts·RK3-steps/s",
        "cores": cores,
        "kind": "port",
        "sample": f"1 RK3 step of the same state, {dt:.1f} s wall, OpenMP threads={cores}",
    }


def measured_traffic(workload, precision, cls):
    """HBM bytes per launch of kernel class `cls` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary for this workload, or None."""
    import glob

    for path in glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_*.md")):
        for ln in open(path):
            ln = ln.strip()
            if ln.startswith('{"workload"'):
                d = json.loads(ln)
                if d["workload"] == workload and d["precision"] == precision:
                    return d["traffic_bytes_per_launch"].get(cls)
    return None


def copy_probe(nbytes):
    """Device-to-device hipMemcpy of nbytes (a plain copy kernel of known
    size): calibrates the HBM byte counters and gives the copy bandwidth."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # DeviceToDevice
        hip.hipDeviceSynchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    hip.hipFree(a)
    hip.hipFree(b)
    return {"bytes": nbytes, "seconds": best, "GBps_read_plus_write": 2 * nbytes / best / 1e9}


def run_form(p, gy, st, precision, steps, warmup, timing_steps):
    """Time `steps` RK3 steps of one form with the state resident, then the
    per-kernel-class split over `timing_steps` more (outside the timed region)."""
    from diablo3_test_amd.binding import ChannelSolver

    npts = p["nx"] * p["ny"] * p["nz"]
    g = ChannelSolver(p, gy, precision=precision)
    g.set_state(st)
    g.pre_first_step()
    g.synchronize()
    log(f"bench: {precision} state resident, warming up")
    g.step(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.step(steps)
    g.synchronize()
    sec = time.perf_counter() - t0
    ms_per_step = sec / steps * 1e3
    log(f"bench: {precision} {steps} steps in {sec:.3f} s, {ms_per_step:.2f} ms/step")
    g.timing_enable(True)
    g.step(timing_steps)
    kt = g.timing_read()
    g.timing_enable(False)
    g.close()
    total_ms = sum(v[0] for v in kt.values())
    total_bytes = sum(v[1] for v in kt.values())
    kernels = {}
    for name, (ms, by, ct) in kt.items():
        if ct:
            kernels[name] = {
                "ms_per_step": ms / timing_steps,
                "launches_per_step": ct / timing_steps,
                "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None,
                "share": ms / total_ms if total_ms > 0 else None,
            }
    step_bytes = total_bytes / timing_steps
    return {
        "value": npts * steps / sec / 1e6,
        "ms_per_step": ms_per_step,
        "kt": kt,
        "kernels": kernels,
        "step_hbm_frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "algorithmic_GB_per_step": step_bytes / 1e9,
    }


def run_adjoint_form(p, gy, st, steps, warmup, timing_steps):
    """The differentiation form (one full untimed run first, whatever `warmup`):
    the gradient of the time-averaged wall shear
    stress over `steps` steps with respect to the initial state -- forward
    run with checkpoints, recomputation and the adjoint substeps.  A step of
    this form is one forward step plus its share of recomputation and one
    adjoint step."""
    from diablo3_test_amd.adjoint import WallShearRun
    from diablo3_test_amd.binding import ChannelSolver

    npts = p["nx"] * p["ny"] * p["nz"]
    g = ChannelSolver(p, gy)
    g.set_state(st)
    g.pre_first_step()
    run = WallShearRun(g, steps, checkpoint_every=steps)
    # one full run first: it allocates the checkpoint slots; slot 0 then holds
    # the initial state, which stays resident for the timed run
    run.gradient(None, fetch=False)
    g.checkpoint(0, save=False)
    g.synchronize()
    t0 = time.perf_counter()
    run.gradient(None, fetch=False)
    g.synchronize()
    sec = time.perf_counter() - t0
    ms_per_step = sec / steps * 1e3
    log(f"bench: adjoint {steps} steps in {sec:.3f} s, {ms_per_step:.2f} ms/step")
    g.checkpoint(0, save=False)
    g.timing_enable(True)
    WallShearRun(g, timing_steps, checkpoint_every=timing_steps).gradient(None, fetch=False)
    kt = g.timing_read()
    g.timing_enable(False)
    g.close()
    kernels = {}
    total_bytes = 0.0
    for name, (ms, by, ct) in kt.items():
        if ct:
            total_bytes += by
            kernels[name] = {"ms_per_step": ms / timing_steps, "launches_per_step": ct / timing_steps,
                             "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None}
    step_bytes = total_bytes / timing_steps
    return {
        "value": npts * steps / sec / 1e6,
        "ms_per_step": ms_per_step,
        "kt": kt,
        "kernels": kernels,
        "step_hbm_frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "algorithmic_GB_per_step": step_bytes / 1e9,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dns_512x257x512")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="the form whose throughput is `value`")
    ap.add_argument("--forms", default="fp64,mixed,adjoint",
                    help="forms also measured and reported under `forms` (comma list, or 'none')")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-steps", type=int, default=2)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a device-to-device copy of this size first")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a build", file=sys.stderr)
        sys.exit(2)

    from diablo3_test_amd import workloads

    probe = copy_probe(args.copy_probe_mib << 20) if args.copy_probe_mib else None
    if probe:
        log(f"bench: copy probe {probe}")

    p, gy, ic, seed, amp = workloads.workload(args.workload)
    npts = p["nx"] * p["ny"] * p["nz"]
    log(f"bench: {args.workload} {p['nx']}x{p['ny']}x{p['nz']}")
    st = workloads.initial_state(p, gy, ic, seed, amp)
    main_form = run_form(p, gy, st, args.precision, args.steps, args.warmup, args.timing_steps)
    forms = {args.precision: main_form}
    extra = [] if args.forms == "none" else [f for f in args.forms.split(",") if f and f != args.precision]
    for f in extra:
        if f == "adjoint":
            if p["use_les"]:
                log("bench: the adjoint covers the DNS substep; skipped for an LES workload")
                continue
            forms[f] = run_adjoint_form(p, gy, st, args.steps, args.warmup, args.timing_steps)
        else:
            forms[f] = run_form(p, gy, st, f, args.steps, args.warmup, args.timing_steps)

    kt = main_form["kt"]
    dname, (dms, dbytes, dct) = max(kt.items(), key=lambda kv: kv[1][0])
    achieved = dbytes / (dms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": measured_traffic(args.workload, args.precision, dname),
        "kernel": dname,
        "bytes_per_launch": dbytes / dct,
        "avg_launch_ms": dms / dct,
    }
    base = forms.get("fp64", main_form)
    line = {
        "metric": METRIC,
        "value": main_form["value"],
        "unit": "Mpts·RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_form["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "grid": f"{p['nx']}x{p['ny']}x{p['nz']}",
            "physics": "DNS" if not p["use_les"] else "LES",
            "precision": args.precision,
            "re": p["re"],
            "n_th": p["n_th"],
        },
        "roofline": roofline,
        "step_hbm_frac": main_form["step_hbm_frac"],
        "algorithmic_GB_per_step": main_form["algorithmic_GB_per_step"],
        "kernels": main_form["kernels"],
        "forms": {
            name: {
                "value": f["value"],
                "ms_per_step": f["ms_per_step"],
                "vs_fp64": f["value"] / base["value"],
                "step_hbm_frac": f["step_hbm_frac"],
                "algorithmic_GB_per_step": f["algorithmic_GB_per_step"],
                "kernels_ms_per_step": {k: v["ms_per_step"] for k, v in f["kernels"].items()},
            }
            for name, f in forms.items()
        },
    }
    if probe:
        line["copy_probe"] = probe
    if not args.no_cpu_baseline:
        log("bench: timing the CPU oracle on one step of the same state")
        line["cpu_baseline"] = cpu_baseline(p, gy, st, npts)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

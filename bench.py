"""Throughput of the channel RK3/CN step on one MI355X.

A step is one pass of the driver loop (for/diablo.f90:98-108): three
rk_chan_1 substeps over the whole field, which stays resident in HBM.
The metric is Mpts.RK3-steps/s = Nx*Ny*Nz * steps / seconds / 1e6.

Beside the GPU number the line carries
* roofline: the dominant kernel class's algorithmic HBM bytes over its
  kernel time, timed live with HIP events on the library's stream;
* cpu_baseline: the CPU oracle (C++ FP64 restatement, OpenMP) timed on the
  host cores on a bounded sample of the same workload (the median of three
  single steps);
* forms: the other forms of the metric -- mixed precision, the adjoint
  (checkpoints every 10 steps, and with one checkpoint), the AMD-Verstappen
  LES in FP64 and mixed -- each with its per-kernel-class split, its
  algorithmic bytes per step, the same without the two-pass transforms'
  x-spectra intermediate, and the reference algorithm's one-pass
  compulsory traffic (model_GB_per_step) over its step time.
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


def host_cpu():
    """lscpu's model name and physical core count of this host."""
    import subprocess

    model, cores_per_socket, sockets = None, None, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            v = v.strip()
            if k.strip() == "Model name":
                model = v
            elif k.strip() == "Core(s) per socket":
                cores_per_socket = int(v)
            elif k.strip() == "Socket(s)":
                sockets = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    phys = cores_per_socket * sockets if cores_per_socket and sockets else None
    return model, phys


def cpu_baseline(p, gy, st, npts, nsteps=3):
    """RK3 steps (three rk_chan_1 substeps each) of the same state on the CPU
    oracle, timed one by one; the median step gives the rate."""
    import ctypes
    import statistics

    from diablo3_test_amd.binding import ChannelSolver, _declare

    lib = _declare(ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so")), "orc_", with_precision=False)
    o = ChannelSolver(p, gy, lib=lib, prefix="orc_")
    o.set_state(st)
    o.pre_first_step()
    times = []
    for i in range(nsteps):
        t0 = time.perf_counter()
        o.step(1)
        times.append(time.perf_counter() - t0)
        log(f"bench: CPU oracle step {i + 1}/{nsteps}: {times[-1]:.2f} s")
    o.close()
    dt = statistics.median(times)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model, phys = host_cpu()
    return {
        "value": npts / dt / 1e6,
        "unit": "Mpts·RK3-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"median of {nsteps} single RK3 steps of the same state "
                   f"({', '.join(f'{t:.2f}' for t in times)} s), 1 process x {threads} OpenMP threads"),
        "cpu_model": model,
        "physical_cores": phys,
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


def model_bytes_per_step(p, form):
    """The one-pass compulsory-traffic model of a step of the reference's
    algorithm (23 two-dimensional transforms per DNS substep, 59 with the
    AMD-Verstappen LES, each reading its input once and writing its output
    once; real CN solves of 4 fields; projection; RHS assembly fused into the
    transforms; the products' extra physical reads; LES pointwise work), in
    bytes: P = Nx Nz (Ny+2) physical points, S = (Nx/2)(2 Nz/3 + 1)(Ny+2)
    spectral modes, r bytes per real.  None for the adjoint form."""
    if form.startswith("adjoint"):
        return None
    r = 4.0 if form.endswith("mixed") else 8.0
    les = form.startswith("les")
    P = p["nx"] * p["nz"] * (p["ny"] + 2)
    S = (p["nx"] // 2) * (2 * (p["nz"] // 3) + 1) * (p["ny"] + 2)
    nfft = 59 if les else 23
    sub = nfft * (r * P + 2 * r * S)
    sub += 4 * ((r + 16) * P + (8 * P if les else 0))
    sub += 128 * S + 17 * 2 * r * S + 8 * r * P
    if les:
        sub += 30 * r * P + 5 * 2 * r * S
    return 3 * sub


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
    tb = g.timing_read_intermediate()
    g.timing_enable(False)
    g.close()
    return summarize(npts, steps, sec, kt, tb, timing_steps)


def summarize(npts, steps, sec, kt, tb, timing_steps):
    """Throughput and the per-kernel-class split of a timed run."""
    ms_per_step = sec / steps * 1e3
    total_ms = sum(v[0] for v in kt.values())
    total_bytes = sum(v[1] for v in kt.values())
    total_tbytes = sum(tb.values())
    kernels = {}
    for name, (ms, by, ct) in kt.items():
        if ct:
            kernels[name] = {
                "ms_per_step": ms / timing_steps,
                "launches_per_step": ct / timing_steps,
                "GBps": by / (ms * 1e-3) / 1e9 if ms > 0 else None,
                # without the x-spectra intermediate of the two-pass transforms
                "GBps_one_pass": (by - tb[name]) / (ms * 1e-3) / 1e9 if ms > 0 else None,
                "share": ms / total_ms if total_ms > 0 else None,
            }
    step_bytes = total_bytes / timing_steps
    return {
        "value": npts * steps / sec / 1e6,
        "ms_per_step": ms_per_step,
        "kt": kt,
        "tb": tb,
        "kernels": kernels,
        "step_hbm_frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "algorithmic_GB_per_step": step_bytes / 1e9,
        "one_pass_GB_per_step": (total_bytes - total_tbytes) / timing_steps / 1e9,
    }


def run_adjoint_form(p, gy, st, steps, warmup, timing_steps, checkpoint_every):
    """The differentiation form (one full untimed run first, whatever
    `warmup`): the gradient of the time-averaged wall shear stress over
    `steps` steps with respect to the initial state -- the forward run saving
    a checkpoint every `checkpoint_every` steps, then per segment (last
    first) the recomputation from its checkpoint and the adjoint substeps.
    The last segment is never recomputed: the forward run keeps its substep
    inputs.  A step of this form is one forward step, its share of
    recomputation and one adjoint step."""
    from diablo3_test_amd.adjoint import WallShearRun
    from diablo3_test_amd.binding import ChannelSolver

    npts = p["nx"] * p["ny"] * p["nz"]
    g = ChannelSolver(p, gy)
    g.set_state(st)
    g.pre_first_step()
    run = WallShearRun(g, steps, checkpoint_every=checkpoint_every)
    # one full run first: it allocates the checkpoint slots; slot 0 then holds
    # the initial state, which stays resident for the timed run
    run.gradient(None, fetch=False)
    g.checkpoint(0, save=False)
    g.synchronize()
    t0 = time.perf_counter()
    run.gradient(None, fetch=False)
    g.synchronize()
    sec = time.perf_counter() - t0
    log(f"bench: adjoint (checkpoints every {checkpoint_every}) {steps} steps in {sec:.3f} s, "
        f"{sec / steps * 1e3:.2f} ms/step")
    g.checkpoint(0, save=False)
    g.timing_enable(True)
    WallShearRun(g, timing_steps, checkpoint_every=min(checkpoint_every, timing_steps)).gradient(None, fetch=False)
    kt = g.timing_read()
    tb = g.timing_read_intermediate()
    g.timing_enable(False)
    g.close()
    out = summarize(npts, steps, sec, kt, tb, timing_steps)
    out["checkpoint_every"] = checkpoint_every
    out["steps"] = steps
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dns_512x257x512")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="the form whose throughput is `value`")
    # adjoint_one_segment runs before adjoint: run after it in the same
    # process its wall time reads ~8 ms/step higher with unchanged kernel time
    ap.add_argument("--forms", default="fp64,mixed,adjoint_one_segment,adjoint,les,les_mixed",
                    help="forms also measured and reported under `forms` (comma list, or 'none'): "
                         "fp64, mixed, adjoint (50 steps, checkpoints every 10), adjoint_one_segment "
                         "(--steps steps, one checkpoint: no recomputation), les and les_mixed (AMD-Verstappen LES "
                         "on the LES workload of the same grid)")
    ap.add_argument("--les-workload", default="cfg3_les_512x257x512")
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
    les_case = None
    for f in extra:
        if f.startswith("adjoint"):
            if p["use_les"]:
                log("bench: the adjoint form is timed on the DNS workload; skipped for an LES workload")
                continue
            if f == "adjoint":  # 50 steps with a checkpoint every 10 (the recomputing cadence)
                forms[f] = run_adjoint_form(p, gy, st, max(args.steps, 50), args.warmup, args.timing_steps, 10)
            else:
                forms[f] = run_adjoint_form(p, gy, st, args.steps, args.warmup, args.timing_steps, args.steps)
        elif f in ("les", "les_mixed"):
            if les_case is None:
                lp, lgy, lic, lseed, lamp = workloads.workload(args.les_workload)
                log(f"bench: LES forms on {args.les_workload}")
                les_case = (lp, lgy, workloads.initial_state(lp, lgy, lic, lseed, lamp))
            lp, lgy, lst = les_case
            forms[f] = run_form(lp, lgy, lst, "mixed" if f == "les_mixed" else "fp64", args.steps, args.warmup,
                                args.timing_steps)
            forms[f]["workload"] = args.les_workload
        elif f in ("fp64", "mixed"):
            forms[f] = run_form(p, gy, st, f, args.steps, args.warmup, args.timing_steps)
        else:
            raise SystemExit(f"unknown form {f}")

    kt = main_form["kt"]
    dname, (dms, dbytes, dct) = max(kt.items(), key=lambda kv: kv[1][0])
    achieved = dbytes / (dms * 1e-3) / 1e9
    one_pass = (dbytes - main_form["tb"][dname]) / (dms * 1e-3) / 1e9
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
        # the same kernel's bytes without its x-spectra intermediate (what a
        # one-pass 2-D transform would move) over the same time
        "achieved_one_pass": one_pass,
        "frac_one_pass": one_pass / HBM_PEAK_GBS,
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
        "forms": {},
    }
    for name, f in forms.items():
        fp = les_case[0] if f.get("workload", args.workload) != args.workload else p
        mb = model_bytes_per_step(fp, name)
        line["forms"][name] = {
            "value": f["value"],
            "ms_per_step": f["ms_per_step"],
            "vs_fp64": f["value"] / base["value"],
            "workload": f.get("workload", args.workload),
            "step_hbm_frac": f["step_hbm_frac"],
            "algorithmic_GB_per_step": f["algorithmic_GB_per_step"],
            "one_pass_GB_per_step": f["one_pass_GB_per_step"],
            # the reference algorithm's one-pass compulsory traffic over this
            # form's step time
            "model_GB_per_step": mb / 1e9 if mb else None,
            "model_hbm_frac": mb / (f["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS if mb else None,
            "kernels": {k: {kk: v[kk] for kk in ("ms_per_step", "GBps", "GBps_one_pass")}
                        for k, v in f["kernels"].items()},
        }
        if "checkpoint_every" in f:
            line["forms"][name]["checkpoint_every"] = f["checkpoint_every"]
            line["forms"][name]["steps"] = f["steps"]
    if "les" in forms and "les_mixed" in forms:
        line["forms"]["les_mixed"]["vs_les"] = forms["les_mixed"]["value"] / forms["les"]["value"]
    if probe:
        line["copy_probe"] = probe
    if not args.no_cpu_baseline:
        log("bench: timing the CPU oracle on three steps of the same state")
        line["cpu_baseline"] = cpu_baseline(p, gy, st, npts)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

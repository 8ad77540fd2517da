#!/usr/bin/env python3
"""Throughput of the NSW6 microphysics step (mp_driver) on one MI355X.

Workload: 10 regions x 16,900 columns = 169,000 columns x 96 levels, dt = 15 s,
consecutive steps on a device-resident state (seeded synthetic NICAM columns).
Prints ONE JSON line: FP64 column-steps/s (the headline value) with the roofline of
the step measured with HIP events on the launch stream; the mixed-precision form's
throughput and its per-field error against the FP64 form; the reverse mode's cost
relative to the forward step at 16,900 x 96; and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

# HBM bytes a step must move at least, per column (96 levels):
# reads 22 fields + writes 27 fields per level (8 B each) + 15 per-column values.
ALGO_BYTES_PER_COLUMN = 96 * (22 + 27) * 8 + 15 * 8          # 37,752 B
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="form of the headline measurement (the other form is reported beside it "
                         "unless --only)")
    ap.add_argument("--only", action="store_true",
                    help="measure the headline form only (no second form, no adjoint; for profiling)")
    ap.add_argument("--adjoint-reps", type=int, default=10)
    ap.add_argument("--rollout", action="store_true",
                    help="instead of the step benchmark: the long rollout (40 regions x 16,900 columns x 96 "
                         "levels, 20 mixed-precision steps, checkpoints every 5 steps, FP64 reverse mode over "
                         "all steps)")
    ap.add_argument("--workload", default=None)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--counters", default=None,
                    help="text file of per-launch counters of the workload, measured with rocprofv3 --pmc "
                         "(default profiles/counters_<workload>.txt): HBM bytes per step (FETCH_SIZE / "
                         "WRITE_SIZE) and the rate kernel's VALU issue cycles and FP64 FLOPs; used only when "
                         "its kernel_sources_sha16 matches the sources of this tree")
    ap.add_argument("--cpu-leg", type=int, default=0, metavar="COLUMNS", help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the measurement, run 3 plain device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE/WRITE_SIZE counters of a profiled run against known bytes)")
    return ap.parse_args()


# MI355X: 256 CUs x 4 SIMDs; a wave64 VALU instruction occupies its SIMD for 4 cycles
# (FP64 FMA/ADD/MUL at full rate), so the issue peak is 1024 x 2.4 GHz SIMD cycles/s and
# the FP64 vector peak 256 x 128 FLOP/clk x 2.4 GHz = 78.6 TFLOP/s.
SIMDS = 1024
PEAK_CLOCK_HZ = 2.4e9
FP64_PEAK_TFLOPS = 78.6

# The sources whose code the per-launch counters describe: the step kernels, their
# compile flags and the synthetic-input generator.  A counters file records their
# digest; after any change to them the stored counters no longer describe the build.
COUNTED_SOURCES = [os.path.join("nicam_kokkos_microphysics_amd", "csrc", n) for n in (
    "nsw6_kernels.hip", "nsw6_rates_kernel.h", "nsw6_rates_mixed.hip", "nsw6_finish.hip", "nsw6_device.h",
    "nsw6_step.h", "nsw6_columns.h", "nsw6_internal.h", "nsw6_setup.h", "nsw6_tables.h", "nsw6_math.h", "nsw6_synth.c")] + \
    [os.path.join("nicam_kokkos_microphysics_amd", "build_lib.py")]


def kernel_sources_sha16():
    import hashlib
    h = hashlib.sha256()
    for rel in COUNTED_SOURCES:
        h.update(rel.encode())
        with open(os.path.join(REPO, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def read_counters(path):
    """'key: value' lines of a counters file -> dict of floats (plus the digest as a
    string); None when the file is missing or was measured on other sources."""
    if not os.path.exists(path):
        return None
    out = {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("#") or ":" not in line:
                continue
            k, v = (x.strip() for x in line.split(":", 1))
            out[k] = v if k == "kernel_sources_sha16" else float(v)
    return out


def host_cores():
    """Physical cores this process may run on: distinct (socket, core) pairs of
    `lscpu -p` among the CPUs of its affinity mask, capped by the cgroup CPU quota.
    Returns (cores, physical cores lscpu lists for the whole host)."""
    import math
    import subprocess
    allowed = os.sched_getaffinity(0)
    pairs, every = set(), set()
    try:
        txt = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
        for line in txt.splitlines():
            if line.startswith("#"):
                continue
            cpu, core, sock = (int(x) for x in line.split(","))
            every.add((sock, core))
            if cpu in allowed:
                pairs.add((sock, core))
    except (OSError, subprocess.CalledProcessError, ValueError):
        pairs = every = {(0, c) for c in allowed}
    cores = len(pairs)
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            quota, period = fh.read().split()
        if quota != "max":
            cores = min(cores, max(1, math.floor(int(quota) / int(period))))
    except (OSError, ValueError):
        pass
    return cores, len(every)


def mixed_error(w, steps):
    """Per-field error and column budget drift of the mixed form against the FP64 form
    (itself held to the oracle within 1e-10 by the parity tests), on region 0 of the
    workload: the error of one mixed step from each state of a `steps`-step FP64
    trajectory (worst over the steps), and the divergence of a free-running mixed
    trajectory after `steps` steps."""
    import numpy as np
    import torch
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.check import budget_drift, field_errors
    from nicam_kokkos_microphysics_amd.driver import NSW6, grid_setup
    gz, gzh, st = F.workload_state(w, lall=1)
    dgz = np.ctypeslib.as_array(grid_setup(gz, gzh).dgz)
    keys = [k for k in F.OUTPUTS if k not in ("rhogq", "q")] + \
        [f"rhogq:{t}" for t in F.TRACERS] + [f"q:{t}" for t in F.TRACERS]
    a, b = F.to_device(st), F.to_device(st)
    d64 = NSW6(gz, gzh, max_columns=w.ijdim, precision=N.NSW6_FP64)
    dmx = NSW6(gz, gzh, max_columns=w.ijdim, precision=N.NSW6_MIXED)
    one_step = {k: 0.0 for k in keys}
    one_drift = {}
    for _ in range(steps):
        c = F.copy_state(a)
        d64.step(a, w.dt)
        dmx.step(c, w.dt)
        dmx.step(b, w.dt)
        torch.cuda.synchronize()
        ha, hc = F.to_host(a), F.to_host(c)
        for k, v in field_errors(hc, ha, keys).items():
            one_step[k] = max(one_step[k], v[2])
        for k, v in budget_drift(hc, ha, dgz).items():
            one_drift[k] = max(one_drift.get(k, 0.0), v)
        del c
    torch.cuda.synchronize()
    d64.close()
    dmx.close()
    ha, hb = F.to_host(a), F.to_host(b)
    free = {k: v[2] for k, v in field_errors(hb, ha, keys).items()}
    worst1 = max(one_step.items(), key=lambda kv: kv[1])
    worstf = max(free.items(), key=lambda kv: kv[1])
    return {"max_field_error": worst1[1], "worst_field": worst1[0], "per_field": one_step,
            "budget_drift": one_drift,
            "free_running": {"max_field_error": worstf[1], "worst_field": worstf[0],
                             "budget_drift": budget_drift(hb, ha, dgz)},
            "sample": f"region 0 ({w.ijdim} columns): worst one-step error over a {steps}-step FP64 "
                      f"trajectory, and a {steps}-step free-running mixed trajectory; error / max-abs "
                      f"of the FP64 field"}


def cpu_leg(columns: int, reps: int):
    """Child process of cpu_baseline: the CPU oracle (an OpenMP FP64 port of the
    reference path) on the first `columns` columns of the FP64 workload (whole regions
    of 16,900), 1 warm-up and the median of `reps` timed mp_driver calls."""
    from nicam_kokkos_microphysics_amd import fields as F
    from oracle import oracle as O
    w = F.WORKLOADS["nsw6_169000x96_fp64"]
    gz, gzh, st = F.workload_state(w, lall=columns // w.ijdim)
    P = O.params()
    O.mp_driver(st, gz, gzh, w.dt, P=P)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.mp_driver(st, gz, gzh, w.dt, P=P)
        times.append(time.perf_counter() - t0)
    print(json.dumps({"median_s": statistics.median(times), "threads": O.num_threads()}))


def cpu_baseline(reps: int):
    """The CPU oracle timed on this host's physical cores (one OpenMP thread per core,
    OMP_PROC_BIND=close, OMP_PLACES=cores), in a child process so that the OpenMP
    runtime starts with that binding: the whole 169,000-column workload, and one region
    of 16,900 columns beside it."""
    import subprocess
    cores, host_total = host_cores()
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    res = {}
    for cols, n in ((169000, max(1, reps // 2)), (16900, reps)):
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", str(cols), "--cpu-reps", str(n)],
                           env=env, capture_output=True, text=True, timeout=900)
        if p.returncode != 0:
            raise RuntimeError(f"CPU baseline at {cols} columns failed:\n{p.stderr[-2000:]}")
        res[cols] = (json.loads(p.stdout.strip().splitlines()[-1]), n)
    (big, nb), (small, ns) = res[169000], res[16900]
    return {"value": 169000 / big["median_s"], "unit": "column-steps/s", "cores": big["threads"], "kind": "port",
            "sample": f"10 regions x 16,900 columns x 96 levels (the workload), dt=15 s, median of {nb} steps after "
                      f"1 warm-up ({big['median_s'] * 1e3:.0f} ms/step); OMP_NUM_THREADS={cores} (physical cores "
                      f"available to the process; lscpu lists {host_total} on the host), OMP_PROC_BIND=close, "
                      f"OMP_PLACES=cores",
            "region_16900": {"value": 16900 / small["median_s"], "unit": "column-steps/s",
                             "ms_per_step": small["median_s"] * 1e3, "steps": ns, "cores": small["threads"]}}


def time_steps(drv, dev, dt, steps, warmup, stream):
    """`warmup` untimed steps, then `steps` timed steps bracketed by device
    synchronisation; also the HIP-event durations of the step and its kernels."""
    import torch
    for _ in range(warmup):
        drv.step(dev, dt, stream=stream, check=False)
    drv.check()
    torch.cuda.synchronize()
    drv.timing(True)
    drv.timing_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        drv.step(dev, dt, stream=stream, check=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rep = drv.check()
    step_ms, nev = drv.timing_read(0)
    kernel_ms = {name: drv.timing_read(i)[0] for i, name in
                 ((1, "rates"), (2, "extremes+bounds"), (3, "flux"), (8, "flux_general"), (4, "finish"))}
    drv.timing(False)
    return wall, step_ms, nev, kernel_ms, rep


def other_form(prec_other, steps, warmup):
    """Throughput of the other precision form on its own workload (169,000 x 96)."""
    import torch
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.driver import NSW6
    wname = "nsw6_169000x96_mixed" if prec_other == N.NSW6_MIXED else "nsw6_169000x96_fp64"
    w = F.WORKLOADS[wname]
    ncol = w.lall * w.ijdim
    gz, gzh, st = F.workload_state(w)
    dev = F.to_device(st)
    del st
    drv = NSW6(gz, gzh, max_columns=ncol, precision=prec_other)
    wall, step_ms, _, kernel_ms, rep = time_steps(drv, dev, w.dt, steps, warmup, torch.cuda.current_stream())
    drv.close()
    del dev
    torch.cuda.empty_cache()
    return {"workload": wname, "value": ncol * steps / wall, "unit": "column-steps/s",
            "ms_per_step": wall * 1e3 / steps, "event_ms": step_ms, "kernels_ms": kernel_ms,
            "general_regions": rep.n_general_regions}


def reference_vgrid():
    """The reference's 96-level vertical grid (data/vgrid/GRD_gz.dat, GRD_gzh.dat: 72 m
    bottom layers), kept beside the tests as a fixture."""
    import numpy as np
    g = os.path.join(REPO, "tests", "golden")
    return (np.fromfile(os.path.join(g, "GRD_gz.dat"), dtype="<f8"),
            np.fromfile(os.path.join(g, "GRD_gzh.dat"), dtype="<f8"))


def reference_vgrid_step():
    """The FP64 step at 169,000 x 96 on the reference vgrid at dt = 30 s with deep
    convection (10 % towers, 1 % rain shafts): 1 warm-up and 5 timed steps.  At
    dt = 15 s no departure on that grid goes beyond the next cell, even under 40 g/kg
    of rain; at 30 s rain falls through two layers and every region takes the general
    wall fluxes.  Reports how many regions did, and the general flux kernel's share."""
    import torch
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.driver import NSW6
    w = F.WORKLOADS["nsw6_169000x96_refgrid_dt30"]
    steps, warmup = w.steps - 1, 1
    ncol = w.lall * w.ijdim
    gz, gzh = reference_vgrid()
    gz, gzh, st = F.workload_state(w, gz=gz, gzh=gzh)
    dev = F.to_device(st)
    del st
    drv = NSW6(gz, gzh, max_columns=ncol, precision=N.NSW6_FP64)
    wall, step_ms, _, kernel_ms, rep = time_steps(drv, dev, w.dt, steps, warmup, torch.cuda.current_stream())
    drv.close()
    del dev
    torch.cuda.empty_cache()
    return {"workload": w.name, "value": ncol * steps / wall, "unit": "column-steps/s",
            "ms_per_step": wall * 1e3 / steps, "event_ms": step_ms, "kernels_ms": kernel_ms,
            "general_regions": rep.n_general_regions, "regions": w.lall,
            "general_flux_share": kernel_ms["flux_general"] / step_ms if step_ms > 0 else None}


def adjoint_cost(reps):
    """Reverse mode (nsw6_mp_driver_adj) against the FP64 forward step on the adjoint
    workload (1 region x 16,900 columns x 96 levels): HIP-event durations on the launch
    stream, `reps` calls of each after one warm-up.  The cotangents are those of the
    adjoint-check loss (2 x surface precipitation, 1/dt on rhoge)."""
    import torch
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.driver import NSW6
    w = F.WORKLOADS["nsw6_16900x96_adjoint"]
    ncol = w.lall * w.ijdim
    gz, gzh, st = F.workload_state(w)
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, max_columns=ncol)
    stream = torch.cuda.current_stream()
    fwd = F.copy_state(dev)
    drv.step(fwd, w.dt, stream=stream)
    cot = {"precip": 2.0 * fwd["precip"], "rhoge": torch.full_like(fwd["rhoge"], 1.0 / w.dt)}
    grad, gcoef = drv.adjoint(dev, cot, w.dt, stream=stream)   # warm-up (allocates the records)
    torch.cuda.synchronize()
    drv.timing(True)
    drv.timing_reset()
    for _ in range(reps):
        s = F.copy_state(dev)
        drv.step(s, w.dt, stream=stream, check=False)
    for _ in range(reps):
        drv.adjoint(dev, cot, w.dt, grad=grad, grad_coef=gcoef, stream=stream, check=False)
    torch.cuda.synchronize()
    rep = drv.check()
    fwd_ms = drv.timing_read(0)[0]
    adj_ms, nadj = drv.timing_read(5)
    adj_fwd_ms = drv.timing_read(6)[0]
    adj_rev_ms = drv.timing_read(7)[0]
    drv.timing(False)
    drv.close()
    return {"workload": w.name, "forward_ms": fwd_ms, "adjoint_ms": adj_ms, "ratio": adj_ms / fwd_ms,
            "adjoint_forward_part_ms": adj_fwd_ms, "adjoint_reverse_passes_ms": adj_rev_ms,
            "calls_timed": nadj, "general_regions": rep.n_general_regions,
            "sample": f"{reps} forward steps and {reps} reverse-mode calls on the same input state, "
                      f"HIP events on the launch stream"}


def adjoint_subcycled_cost(reps, div_num=2):
    """Reverse mode of a sub-cycled call (MP_DIV_NUM = div_num) against the sub-cycled
    forward call on the adjoint workload: CUDA events on the launch stream around
    `reps` calls of each, after one warm-up of each."""
    import torch
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.driver import NSW6
    w = F.WORKLOADS["nsw6_16900x96_adjoint"]
    gz, gzh, st = F.workload_state(w)
    dev = F.to_device(st)
    drv = NSW6(gz, gzh, max_columns=w.lall * w.ijdim)
    drv.set_div_num(div_num)
    stream = torch.cuda.current_stream()
    fwd = F.copy_state(dev)
    drv.step(fwd, w.dt, stream=stream)
    cot = {"precip": 2.0 * fwd["precip"], "rhoge": torch.full_like(fwd["rhoge"], 1.0 / w.dt)}
    grad, gcoef = drv.adjoint(dev, cot, w.dt, stream=stream)   # warm-up (allocates the scratch)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(reps):
        drv.step(fwd, w.dt, stream=stream, check=False)
    ev[1].record(stream)
    for _ in range(reps):
        drv.adjoint(dev, cot, w.dt, grad=grad, grad_coef=gcoef, stream=stream, check=False)
    ev[2].record(stream)
    torch.cuda.synchronize()
    rep = drv.check()
    fwd_ms = ev[0].elapsed_time(ev[1]) / reps
    adj_ms = ev[1].elapsed_time(ev[2]) / reps
    drv.close()
    return {"workload": w.name, "div_num": div_num, "forward_call_ms": fwd_ms, "adjoint_call_ms": adj_ms,
            "ratio": adj_ms / fwd_ms, "calls_timed": reps, "general_regions": rep.n_general_regions}


def subcycling_cost(steps, warmup, div_num=2):
    """MP_DIV_NUM sub-cycling on the FP64 workload (169,000 x 96): wall time of one
    mp_driver call of `div_num` sub-steps of dt / div_num, against div_num plain steps."""
    import torch
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.driver import NSW6
    w = F.WORKLOADS["nsw6_169000x96_fp64"]
    ncol = w.lall * w.ijdim
    gz, gzh, st = F.workload_state(w)
    dev = F.to_device(st)
    del st
    drv = NSW6(gz, gzh, max_columns=ncol, div_num=div_num)
    for _ in range(warmup):
        drv.step(dev, w.dt, check=False)
    drv.check()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        drv.step(dev, w.dt, check=False)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    rep = drv.check()
    drv.close()
    del dev
    torch.cuda.empty_cache()
    return {"workload": w.name, "div_num": div_num, "calls": steps, "ms_per_call": wall * 1e3 / steps,
            "column_steps_per_s": ncol * steps / wall, "general_regions": rep.n_general_regions}


def rollout_bench():
    """The long rollout: forward (mixed precision) and backward (recompute + FP64 reverse
    mode, 10 regions at a time) wall times, and the peak device memory."""
    import torch
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.rollout import Rollout
    w = F.WORKLOADS["nsw6_676000x96_rollout"]
    every = 5
    gz, gzh, st = F.workload_state(w)
    dev = F.to_device(st)
    del st
    rhoge0 = dev["rhoge"].clone()
    ro = Rollout(gz, gzh, w.lall, w.ijdim, precision=N.NSW6_MIXED, chunk_regions=10)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    ckpt, precip = ro.forward(dev, w.dt, w.steps, every)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    L = ro.loss(rhoge0, dev, precip, w.dt)
    grad, gcoef = ro.backward(dev, ckpt, precip, w.dt, w.steps, every)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ncol = w.lall * w.ijdim
    gnorm = float(sum(float((g * g).sum()) for g in grad.values())) ** 0.5
    out = {"metric": "long rollout: column-steps/s forward and backward", "workload": w.name,
           "columns": ncol, "levels": 96, "steps": w.steps, "checkpoint_every": every,
           "forward_s": t1 - t0, "backward_s": t2 - t1,
           "forward_column_steps_per_s": ncol * w.steps / (t1 - t0),
           "backward_column_steps_per_s": ncol * w.steps / (t2 - t1),
           "backward_over_forward": (t2 - t1) / (t1 - t0),
           "loss": L, "grad_norm": gnorm, "grad_coef": [float(x) for x in gcoef.cpu()],
           "peak_device_GiB": torch.cuda.max_memory_allocated() / 2 ** 30,
           "forward_precision": "mixed", "reverse_precision": "f64", "data": "synthetic",
           "note": "backward = per-segment recomputation from the checkpoint + FP64 reverse mode per step, "
                   "10 regions per reverse-mode call; includes the per-call status synchronisation"}
    ro.close()
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the step shards by NICAM region with no "
                                   "communication, and multi-GPU runs are not part of it", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.cpu_leg:
        cpu_leg(args.cpu_leg, args.cpu_reps)
        return
    if args.rollout:
        rollout_bench()
        return
    import torch
    from nicam_kokkos_microphysics_amd import _native as N
    from nicam_kokkos_microphysics_amd import fields as F
    from nicam_kokkos_microphysics_amd.driver import NSW6

    wname = args.workload or ("nsw6_169000x96_mixed" if args.precision == "mixed" else "nsw6_169000x96_fp64")
    w = F.WORKLOADS[wname]
    ncol = w.lall * w.ijdim
    if wname == "nsw6_169000x96_refgrid_dt30":   # its state is built on the reference vgrid
        gz, gzh = reference_vgrid()
        gz, gzh, st = F.workload_state(w, gz=gz, gzh=gzh)
    else:
        gz, gzh, st = F.workload_state(w)
    dev = F.to_device(st)
    del st
    prec = N.NSW6_MIXED if args.precision == "mixed" else N.NSW6_FP64
    drv = NSW6(gz, gzh, max_columns=ncol, precision=prec)
    stream = torch.cuda.current_stream()
    wall, step_ms, nev, kernel_ms, rep = time_steps(drv, dev, w.dt, args.steps, args.warmup, stream)
    ms_per_step = wall * 1e3 / args.steps
    algo_bytes = ALGO_BYTES_PER_COLUMN * ncol
    achieved = algo_bytes / (step_ms * 1e-3) / 1e9 if step_ms > 0 else 0.0
    cpath = args.counters or os.path.join(REPO, "profiles", f"counters_{wname}.txt")
    counters = read_counters(cpath)
    sha = kernel_sources_sha16()
    stale = counters is not None and counters.get("kernel_sources_sha16") != sha
    if stale:
        counters = None
    traffic = counters.get("bytes_per_step") if counters else None
    valu = None
    rates_ms = kernel_ms["rates"]
    if counters and "k_rates_valu_cycles" in counters and rates_ms > 0:
        cyc, flop = counters["k_rates_valu_cycles"], counters["k_rates_fp64_flop"]
        tflops = flop / (rates_ms * 1e-3) / 1e12
        valu = {"kernel": "k_rates", "bound": "valu", "launch_ms": rates_ms,
                "issue_frac": cyc / (SIMDS * PEAK_CLOCK_HZ * rates_ms * 1e-3),
                "fp64_tflops": tflops, "fp64_peak_tflops": FP64_PEAK_TFLOPS, "fp64_frac": tflops / FP64_PEAK_TFLOPS,
                "valu_cycles_per_launch": cyc, "fp64_flop_per_launch": flop}

    out = {
        "metric": "column-steps/s at 169k x 96, FP64 vs mixed at stated error; adjoint/forward ratio",
        "form": "fp64" if prec == N.NSW6_FP64 else "mixed",
        "value": ncol * args.steps / wall,
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == N.NSW6_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": wname, "regions": w.lall, "columns_per_region": w.ijdim, "levels": 96,
                   "columns": ncol, "dt_s": w.dt, "tracers": 6},
        "roofline": {"bound": "hbm", "kernel": "nsw6 step (k_rates + k_extremes + k_bounds + k_flux + k_flux_general + k_finish)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": algo_bytes, "event_ms": step_ms, "launches_timed": nev,
                     "valu_frac": valu["issue_frac"] if valu else None, "valu": valu,
                     "counters": os.path.relpath(cpath, REPO) if counters else None,
                     "counters_stale": stale, "kernel_sources_sha16": sha},
        "kernels_ms": kernel_ms,
        "general_regions": rep.n_general_regions,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_reps)
    drv.close()
    del dev
    if not args.only:
        other = N.NSW6_FP64 if prec == N.NSW6_MIXED else N.NSW6_MIXED
        out["fp64" if other == N.NSW6_FP64 else "mixed"] = other_form(other, args.steps, args.warmup)
    if prec == N.NSW6_MIXED:
        out["error_vs_fp64"] = mixed_error(w, w.steps)
    elif not args.only:
        out["mixed"]["error_vs_fp64"] = mixed_error(F.WORKLOADS["nsw6_169000x96_mixed"], w.steps)
    if not args.only:
        out["reference_vgrid_dt30"] = reference_vgrid_step()
        out["adjoint"] = adjoint_cost(args.adjoint_reps)
        out["adjoint_subcycled"] = adjoint_subcycled_cost(args.adjoint_reps)
        out["subcycling"] = subcycling_cost(max(2, args.steps // 2), 1)
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = 2 * n * 8
    print(json.dumps(out))


if __name__ == "__main__":
    main()

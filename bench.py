"""Benchmark of the elastic time-step hot path on one MI355X.

Workload (the mixed-precision elastic configuration): 8192x8192 cells,
dx = 5 m, seeded 20-layer model (Vp 1.5-4 km/s per layer, 5% correlated
perturbation, Vs = Vp/1.73, rho = 310 Vp^0.25), 15 Hz Ricker vertical force at
100 m depth, free-surface SAT on top, characteristic SAT on the other sides,
256 surface receivers.  Fields stored in BF16 with the Kahan residual of each
store kept as an 8-bit code (compensated time-stage sums), FP32 arithmetic,
material coefficients stored as 8-bit codes of a per-row affine map; the FP32
and FP64 comparison runs use the same quantized material, and the line reports
how far that material is from the exact FP64 one (material_max_rel_err).

A step is one full leapfrog step (velocity and stress half steps, SAT,
compensated update, source, receivers, FP64 energy) of all 5 fields.
value = cell updates per second over the K timed steps, inputs resident in
HBM.  The same run times the FP32-storage solver on the same workload,
measures the BF16 error after the configuration's 2000 steps against FP64 runs
on the same quantized material and on the exact material, times and checks
other storage variants (by default 16-bit coefficients, uncompensated BF16,
FP16 with 8-bit residual codes, and BF16 / FP16 whose stencils also read the
residual codes with 16-bit coefficient codes), names those within the 1e-4
target (accuracy.meets_target), reports the library's streaming-copy
bandwidth beside the HBM peak, and times the CPU oracle and the reference's
own CPU solvers on bounded samples.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size M]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tpu-wave_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "grid-pt updates/s & % HBM peak; BF16-store vs FP32 at ≤1e-4 rel L2 vs FP64"


def comp_bytes(s, cpst):
    """Bytes of the stored compensation per value: a storage-type companion
    array, one byte for the K8 residual code, none for cpst 0."""
    if cpst in ("K8", "K8S"):
        return 1
    return s if cpst in ("3R", "-3R", "6R", "3C", "-3C", "6C", "K") else 0


def coef_bytes(precision, coef16):
    """Bytes per stored material coefficient: FP64 with FP64 storage, else
    FP32, the 16-bit type (coef16 1), an 8-bit code (coef16 2) or a 16-bit
    code (coef16 3; the per-row offset and step of the codes add 8 bytes per
    row, under 0.01 B per cell)."""
    if precision == 0:
        return 8
    if precision == 1:
        return 4
    return {0: 4, 1: 2, 2: 1, 3: 2}[coef16]


def algorithmic_bytes(precision, cpst, coef16, physics=0):
    """Compulsory HBM bytes per cell update: 2 F s (state r+w) + 2 F c
    (stored compensation r+w) + K p (coefficients read)."""
    F, K = (5, 5) if physics == 0 else (3, 3)
    s = {0: 8, 1: 4, 2: 2, 3: 2}[precision]
    c = comp_bytes(s, cpst)
    return 2 * F * s + 2 * F * c + K * coef_bytes(precision, coef16)


def half_bytes(precision, cpst, coef16, half, physics=0):
    """Compulsory HBM bytes per cell of one half step's own fields: the state
    and stored compensation of the fields it updates (read + write) and its
    coefficients (read).  Velocity (half 0): 2 fields, 2 coefficient arrays;
    stress (half 1): 3 fields and 3 arrays elastic, 1 and 1 acoustic.  The two
    halves sum to algorithmic_bytes."""
    F = 2 if half == 0 else (3 if physics == 0 else 1)
    s = {0: 8, 1: 4, 2: 2, 3: 2}[precision]
    c = comp_bytes(s, cpst)
    return 2 * F * s + 2 * F * c + F * coef_bytes(precision, coef16)


def kernel_source_hash():
    """SHA-256 (first 16 hex digits) of the kernel sources (the HIP kernels and
    launches, and the plan they read): a measured traffic figure is used only
    for the kernels it was measured on."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "tpu-wave_amd", "csrc")
    for name in ("tw_internal.h", "tw_kernels.hip"):
        with open(os.path.join(d, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(key, kernel):
    """HBM bytes per launch of `kernel` ('step' for a whole step) measured with
    rocprofv3 (FETCH_SIZE + WRITE_SIZE, calibrated by a copy kernel), from
    profiles/traffic.txt lines 'traffic <workload-key> <kernel-hash> <kernel>
    <bytes>'; None if absent or measured on other kernel sources."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    kh = kernel_source_hash()
    with open(path) as f:
        for line in f:
            t = line.split()
            if len(t) == 5 and t[0] == "traffic" and t[1] == key and t[2] == kh and t[3] == kernel:
                return float(t[4])
    return None


def copy_calibration(torch, nbytes, reps, kernel="torch"):
    """Plain device copy of nbytes, reps times: the known-size kernel that
    scales the FETCH_SIZE / WRITE_SIZE counters, and the attainable
    read+write bandwidth of a streaming copy on this device.  kernel "torch":
    torch's copy; "tw": the library's streaming copy with non-temporal loads
    and stores (the load kind the solver's once-read rows use)."""
    a = torch.ones(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    st = torch.cuda.current_stream()
    if kernel == "tw":
        import twave as T
        L = T.lib()
        run = lambda: L.tw_stream_copy(b.data_ptr(), a.data_ptr(), nbytes, st.cuda_stream)
    else:
        run = lambda: b.copy_(a)
    run()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        run()
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    print(json.dumps({"copy_kernel": kernel, "copy_bytes": nbytes, "reps": reps + 1,
                      "copy_read_plus_write_GBps": 2 * nbytes / t / 1e9}))


def stream_copy_bandwidth(torch, T, nbytes=1 << 30, reps=20):
    """Read + write bandwidth (GB/s) of the library's hand-written streaming
    copy (tw_stream_copy: 16-byte non-temporal loads and stores) of nbytes,
    timed with HIP events on its stream: the attainable rate of a kernel that
    moves every byte once, the ceiling the solver's fraction is read against."""
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    st = torch.cuda.Stream()
    L = T.lib()
    for _ in range(2):
        assert L.tw_stream_copy(b.data_ptr(), a.data_ptr(), nbytes, st.cuda_stream) == 0
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        L.tw_stream_copy(b.data_ptr(), a.data_ptr(), nbytes, st.cuda_stream)
    e1.record(st)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    t = e0.elapsed_time(e1) / 1e3 / reps
    del a, b
    return 2 * nbytes / t / 1e9


def make_solver(T, W, m, precision, cpst, coef16, nt, model_from=None):
    """The elastic workload on a new solver; model_from: take that solver's
    (quantized) coefficients instead of the exact interpolation."""
    cfg = W.elastic_config(m=m, precision=precision, cpst=cpst, coef16=coef16, nt=nt)
    s = T.Solver(cfg)
    W.setup_elastic(s, m, 0.005)
    s.material_err = None
    if model_from is not None:  # same (quantized) coefficients as the low-precision run
        err = 0.0
        for k in (T.C_BINV_VX, T.C_BINV_VY, T.C_MU, T.C_LAMBDA, T.C_TWOMU):
            q = model_from.get_coef(k)
            if precision == T.FP64:  # its own coefficients are the exact FP64 interpolation
                own = s.get_coef(k)
                err = max(err, float(np.max(np.abs(q - own) / np.abs(own))))
                del own
            s.set_coef(k, q)
        if precision == T.FP64:
            s.material_err = err
    return s


PROFILED_STEPS = 100


def timed_steps(torch, s, steps, warmup, profile=False):
    """Device time (s) of `steps` steps on the solver's stream (after `warmup`
    untimed steps) and the wall time; with profile, also the summed durations
    (ms) of the velocity and stress core launches of the first PROFILED_STEPS
    steps of the timed region (HIP events on that stream, four per step: on
    every step they would add about 2% to the step time)."""
    stream = torch.cuda.Stream()
    s.set_stream(stream.cuda_stream)
    if warmup:
        s.step(warmup)
    s.synchronize()
    torch.cuda.synchronize()
    if profile:
        s.profile_begin(min(steps, PROFILED_STEPS))
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    s.step(steps)
    e1.record(stream)
    s.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = s.profile_end() if profile else None
    return e0.elapsed_time(e1) / 1e3, wall, prof


def cpu_threads():
    """Threads a CPU baseline may use: the CPUs this process may run on,
    bounded by OMP_NUM_THREADS when the host sets it (the share of the host's
    cores that comes with the GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cpu_baseline(m, seconds_target=15.0):
    """Oracle (FP32, all host threads) on the same workload, bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import twave as T
    import twave.workloads as W
    threads = cpu_threads()
    O.set_threads(threads)
    cap = 64
    cfg = W.elastic_config(m=m, precision=T.FP32, nt=cap)
    oc = O.make_config(physics=0, mx=m, my=m, dx=cfg.dx, num_dx=cfg.num_dx, den_dx=cfg.den_dx, dt=cfg.dt,
                       bc="FCCC", precision=O.FP32, nt_max=cap)
    o = O.Oracle(oc)
    rho, vp, vs = W.layered_elastic_model(m, 0.005)
    o.set_model(rho, vp, vs, 1, 200)
    o.set_source_ricker(O.GRID_NM, 20, m // 2, 15.0, 1.25 / 15.0)
    o.step(1)
    n, t = 0, 0.0
    while t < seconds_target and n < cap - 2:
        t0 = time.perf_counter()
        o.step(1)
        t += time.perf_counter() - t0
        n += 1
    return {"value": m * m * n / t, "unit": "grid-point updates/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (oracle/tw_oracle.c, FP32, OpenMP) on the same %dx%d elastic workload, "
                      "%d steps (%.1f s) incl. per-step FP64 energy" % (m, m, n, t)}


def _reference_loop_seconds(cmd, timeout_s):
    """Run a reference acoustic binary (it writes its results to ../result)
    in a scratch folder and return the "loop time" it prints, or None."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        work = os.path.join(d, "a")
        os.makedirs(work)
        try:
            r = subprocess.run(cmd, cwd=work, capture_output=True, text=True, timeout=timeout_s,
                               stdin=subprocess.DEVNULL)
        except (subprocess.TimeoutExpired, OSError):
            return None
        if r.returncode != 0:
            return None
        for line in r.stdout.splitlines():
            if line.strip().startswith("loop time"):
                return float(line.split(":")[-1])
    return None


def _grid_dims(n):
    a = int(math.sqrt(n))
    while n % a:
        a -= 1
    return n // a, a


def reference_cpu_samples(threads, timeout_s=120.0):
    """The reference's own acoustic CPU solvers, built from their sources by
    oracle/ref_build.sh, on their 600x600 periodic FP32 case
    (singlecore/src/main_cpp/cpu_acoustic.cpp, mpi/src/main_cpp/cpu_acoustic.cpp):
    the single-core program as an -O3 -DNDEBUG build and with the reference's
    own flags (-O3, asserts on; singlecore/build/makefile:44 adds -march=native,
    left out so the binary runs on any host), and the MPI program on `threads`
    ranks where the image has MPI.  Updates/s from the "loop time" each prints."""
    import shutil
    ref = os.path.join(ROOT, "oracle", "_ref")
    out = {}
    runs = [("singlecore_ndebug", os.path.join(ref, "singlecore_cli_f32"), 3000, 1, None),
            ("singlecore_refflags", os.path.join(ref, "singlecore_cli_f32_refflags"), 500, 1, None)]
    mpiexec = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.access("/opt/conda/bin/mpiexec", os.X_OK)
                                          else None)
    if mpiexec:
        runs.append(("mpi_refflags", os.path.join(ref, "mpi_acoustic_f32_refflags"), 2000, threads, mpiexec))
    for name, exe, nt, np_, launcher in runs:
        if not os.access(exe, os.X_OK):
            out[name] = {"value": None, "note": "not built (oracle/ref_build.sh needs the reference tree%s)" %
                         (" and MPI" if launcher else "")}
            continue
        cmd = [exe, "-Nt", str(nt)]
        if launcher:
            dims = _grid_dims(np_)
            cmd = [launcher, "-n", str(np_), exe, "-Nt", str(nt), "-procs", "%d %d" % dims]
        secs = _reference_loop_seconds(cmd, timeout_s)
        if secs is None or secs <= 0:
            out[name] = {"value": None, "note": "did not run here: %s" % " ".join(os.path.basename(c) for c in cmd)}
            continue
        out[name] = {"value": 600 * 600 * nt / secs, "unit": "grid-point updates/s", "cores": np_,
                     "kind": "reference", "sample": "%s, 600x600 periodic FP32, %d steps, loop %.2f s" %
                     (os.path.basename(exe), nt, secs)}
    if not mpiexec:
        out["mpi_refflags"] = {"value": None, "note": "no MPI launcher on this host"}
    return out


def gpu_reference_case(torch, T, nt=2000):
    """This solver on the reference's own acoustic CPU case (600x600 periodic,
    FP32, dx = 0.008, dt = 1e-4, its stencil and energy forms), for a
    like-for-like ratio with reference_cpu_samples."""
    cfg = T.config(physics=T.ACOUSTIC, mx=600, my=600, dx=(1.0 / (5.0 * 2.5)) / 10, num_dx=0, den_dx=0, dt=1e-4,
                   bc="PPPP", precision=T.FP32, cpst="0", stencil_form=1, energy_form=1, nt_max=nt + 100)
    s = T.Solver(cfg)
    s.set_source_ricker(T.GRID_MM, 200, 200, 5.0, 0.25)
    s.set_receivers([T.VX, T.VY, T.P], [400, 400, 400], [400, 400, 400])
    t, _, _ = timed_steps(torch, s, nt, 100)
    s.close()
    return {"value": 600 * 600 * nt / t, "unit": "grid-point updates/s", "ms_per_step": t / nt * 1e3,
            "sample": "600x600 periodic acoustic FP32 (the reference CPU case), %d steps, energy and 3 receivers "
                      "recorded every step" % nt}


def c4_batch(args):
    """Batched acoustic shots: 64 problems of n x n cells, n uniform in
    [512, 2048] (seeded), dx = 10 m, each with its own smooth random velocity
    (1.5-4.5 km/s, Gaussian-filtered noise, 100 m) and a 10-25 Hz Ricker source
    at a random interior point, characteristic sides, BF16 storage with FP32
    arithmetic; one batched launch per half step, the 1500-step loop in a
    hipGraph, tile height from the autotuner.  value = cell updates/s of the
    whole batch over the graph's device time."""
    import torch
    import twave as T
    import twave.workloads as W
    torch.cuda.set_device(0)
    rng = np.random.default_rng(4)
    nprob, K = args.shots, args.c4_steps
    dx = 0.01
    vmax = 4.5
    sizes = [int(v) for v in rng.integers(512, 2049, nprob)]
    solvers = []
    for i, n in enumerate(sizes):
        rho, vp, _ = W.smooth_acoustic_model(n, dx, seed=1000 + i)
        cfg = T.config(physics=T.ACOUSTIC, mx=n, my=n, dx=dx, num_dx=1, den_dx=100,
                       dt=0.5 * dx / (math.sqrt(2.0) * vmax), bc="CCCC", precision=T.PRECISION[args.precision],
                       cpst="0", coef16=args.c4_coef16, nt_max=K + 64, record_energy=0, stencil_form=1)
        s = T.Solver(cfg)
        s.set_model(rho, vp, vp, 1, 100)
        f0 = float(rng.uniform(10.0, 25.0))
        ix, iy = (int(v) for v in rng.integers(n // 8, n - n // 8, 2))
        s.set_source_ricker(T.GRID_MM, ix, iy, f0, 1.25 / f0)
        solvers.append(s)
    B = T.Batch(solvers)
    rows = B.autotune(20)
    B.step(K)
    ms = B.elapsed_ms()
    cells = sum(n * n for n in sizes)
    value = cells * K / (ms / 1e3)
    prec = T.PRECISION[args.precision]
    Bu = algorithmic_bytes(prec, "0", args.c4_coef16, physics=1)
    out = {"metric": "grid-pt updates/s (batched acoustic shots)", "value": value, "unit": "grid-point updates/s",
           "n_gpus": 1, "steps": K, "warmup": 0, "ms_per_step": ms / K, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "64 acoustic shots, n x n with n uniform in [512, 2048], dx 10 m, smooth random "
                                  "velocity 1.5-4.5 km/s, Ricker 10-25 Hz, CCCC SAT, storage %s, coef16 %d" %
                                  (args.precision, args.c4_coef16),
                      "problems": nprob, "cells": cells, "tile_rows": rows},
           "roofline": {"bound": "hbm", "achieved": value * Bu / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": value * Bu / 1e9 / HBM_PEAK_GBS, "traffic": None, "bytes_per_update": Bu,
                        "kernel": "one step = batched velocity + stress launches of all problems"}}
    print(json.dumps(out))


def c5_adjoint(args):
    """The adjoint-gradient configuration: acoustic 2048^2, 2 km/s background
    plus a seeded circular 10% anomaly, Ricker 12 Hz shots on p under the free
    surface, 512 receivers, 3000 steps (twave.workloads.AdjointShots); the
    FP64 data of --c5-shots shots, then the gradient of each shot's misfit
    (checkpointed forward, backward sweep) timed in the storage of
    --c5-precision from the background model.  value = gradients per second;
    the per-step split comes from rocprofv3 of this command."""
    import torch
    import twave as T
    import twave.workloads as W
    torch.cuda.set_device(0)
    n, nt, k = 2048, 3000, args.c5_shots
    prec = T.PRECISION[args.c5_precision]
    ref = W.AdjointShots(n=n, nt=nt, nshots=16, nrcv=512, precision=T.FP64)
    P = ref if prec == T.FP64 else W.AdjointShots(n=n, nt=nt, nshots=16, nrcv=512, precision=prec, data=ref.data)
    t = []
    for i in range(k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.run(P.m0, i, P.data[i])
        torch.cuda.synchronize()
        t.append(time.perf_counter() - t0)
    if P is not ref:
        P.close()
    ref.close()
    sec = float(np.mean(t))
    out = {"metric": "adjoint gradients per second (acoustic 2048^2, 3000 steps, 512 receivers)",
           "value": 1.0 / sec, "unit": "shot gradients/s", "n_gpus": 1, "steps": nt, "warmup": 0,
           "ms_per_step": sec / nt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": args.c5_precision, "data": "synthetic",
           "config": {"workload": "acoustic 2048x2048, dx 10 m, 2 km/s + 10% circular anomaly, FCCC SAT, "
                                  "Ricker 12 Hz shots on p at row 4, 512 receivers, 3000 steps, sqrt(N) checkpoints",
                      "shots_timed": k, "seconds_per_shot": [round(v, 4) for v in t]}}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--precision", default="bf16")
    ap.add_argument("--cpst", default="K8",
                    help="compensation; K8 = the Kahan residual kept as an 8-bit code per value (the mixed-precision "
                    "config), K = 16-bit residual, 0 = none")
    ap.add_argument("--coef16", type=int, default=2,
                    help="material coefficient storage: 0 FP32, 1 the 16-bit type, 2 one byte per value with a "
                    "per-row affine map (default), 3 two bytes per value with a per-row affine map")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP32, FP64 and variant runs")
    ap.add_argument("--accuracy-steps", type=int, default=2000,
                    help="step count at which every run is compared with FP64 (untimed beyond the timed range; "
                    "the configuration's 2000 steps by default)")
    ap.add_argument("--variants", default="bf16:K8:1,bf16:0:2,fp16:K8:2,bf16:K8S:3,fp16:K8S:3",
                    help="comma-separated storage:cpst[:coef16] triples (coef16 default 1) timed and checked "
                    "against FP64 beside the main run, on the same material")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-profile", action="store_true",
                    help="no per-launch events in the timed region (roofline from the step time)")
    ap.add_argument("--cpu-size", type=int, default=0, help="grid of the CPU sample (default: --size)")
    ap.add_argument("--copy-kernel", default="torch", choices=["torch", "tw"],
                    help="copy of --copy-calib: torch's, or the library's non-temporal streaming copy")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    ap.add_argument("--workload", default="elastic", choices=["elastic", "c4", "c5"],
                    help="elastic: the mixed-precision elastic case (default); c4: batched acoustic shots; "
                    "c5: adjoint gradients of the 2048^2 acoustic shots")
    ap.add_argument("--c5-shots", type=int, default=2, help="c5: shot gradients timed")
    ap.add_argument("--c5-precision", default="fp32", choices=["fp32", "fp64"], help="c5: storage")
    ap.add_argument("--shots", type=int, default=64, help="c4: number of problems")
    ap.add_argument("--c4-steps", type=int, default=1500, help="c4: time steps")
    ap.add_argument("--c4-coef16", type=int, default=1,
                    help="c4: coefficient storage (16-bit by default: the acoustic batch runs no faster with 8-bit)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the library has no multi-GPU path (the slab decomposition "
              "across GPUs is a design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calib:
        copy_calibration(torch, args.copy_calib, 5, args.copy_kernel)
        return
    if args.workload == "c4":
        c4_batch(args)
        return
    if args.workload == "c5":
        c5_adjoint(args)
        return
    import twave as T
    import twave.workloads as W
    torch.cuda.set_device(0)
    m, K, Wn = args.size, args.steps, args.warmup
    nt = K + Wn
    prec = T.PRECISION[args.precision]

    # every run advances to the same step count (the configuration's 2000 steps
    # unless the timed range is longer) before fields and traces are compared
    n_acc = max(args.accuracy_steps, nt)

    s = make_solver(T, W, m, prec, args.cpst, args.coef16, n_acc)
    t_gpu, wall, prof = timed_steps(torch, s, K, Wn, profile=not args.no_profile)
    value = m * m * K / t_gpu
    B = algorithmic_bytes(prec, args.cpst, args.coef16)
    key = "elastic-%d-%s-%s-%d" % (m, args.precision, args.cpst, args.coef16)
    step_ach = B * m * m / (t_gpu / K) / 1e9
    step = {"achieved": step_ach, "frac": step_ach / HBM_PEAK_GBS, "bytes_per_update": B,
            "traffic": measured_traffic(key, "step")}
    if prof:
        # dominant kernel: the core launch of the half step with the larger device time
        ms_v, ms_s, n = prof
        half = 1 if ms_s >= ms_v else 0
        name = ("k_str3", "k_vel3")[half == 0]
        us = (ms_s if half else ms_v) / n * 1e3
        Bh = half_bytes(prec, args.cpst, args.coef16, half)
        achieved = Bh * m * m / (us * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(key, name),
                "kernel": name + " (%s half step, core region)" % ("stress" if half else "velocity"),
                "launch_us": us, "bytes_per_cell": Bh, "other_core_launch_us": (ms_v if half else ms_s) / n * 1e3,
                "traffic_unit": "HBM bytes per launch", "traffic_key": key, "kernel_hash": kernel_source_hash(),
                "step": step}
    else:
        roof = dict(step, bound="hbm", peak=HBM_PEAK_GBS, unit="GB/s",
                    kernel="one step = velocity + stress half-step kernels")
    copy_gbps = stream_copy_bandwidth(torch, T)
    roof["copy_attainable"] = {"GBps": copy_gbps, "frac_of_peak": copy_gbps / HBM_PEAK_GBS,
                               "kernel": "tw_stream_copy, 1 GiB, 16-byte non-temporal loads and stores"}
    roof["frac_of_attainable"] = roof["achieved"] / copy_gbps
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wn,
        "ms_per_step": t_gpu / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec != T.FP64 else "f64",
        "data": "synthetic",
        "config": {"workload": "elastic %dx%d, dx=5 m, 20-layer model, Ricker 15 Hz at 100 m, FCCC SAT, "
                               "256 receivers; storage %s, cpst %s, coef16 %d" %
                               (m, m, args.precision, args.cpst, args.coef16),
                   "cells": m * m, "fields": 5, "storage": args.precision, "cpst": args.cpst,
                   "coef16": args.coef16},
        "roofline": roof,
    }
    if not args.no_compare:
        def finish(sv):
            """Advance to n_acc steps (untimed) and return the compensated
            fields and the receiver traces."""
            sv.step(n_acc - sv.current_step())
            sv.synchronize()
            return {f: sv.get_field_compensated(f) for f in sv.fields}, sv.receivers()

        def rl2(a, b):
            return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

        def errs(run, ref):
            (fa, ra), (fb, rb) = run, ref
            return {"fields": max(rl2(fa[f], fb[f]) for f in fb), "receivers": rl2(ra, rb)}

        main_run = finish(s)
        # FP64 on the same (quantized) material: the error of the state storage alone
        s64 = make_solver(T, W, m, T.FP64, "0", 0, n_acc, model_from=s)
        material_err = s64.material_err
        ref_q = finish(s64)
        s64.close()
        # FP64 on the exact material: the error of storage and coefficient quantization together
        s64 = make_solver(T, W, m, T.FP64, "0", 0, n_acc)
        ref_x = finish(s64)
        s64.close()
        s32 = make_solver(T, W, m, T.FP32, "0", 0, n_acc, model_from=s)
        t32, _, _ = timed_steps(torch, s32, K, Wn)
        out["fp32_value"] = m * m * K / t32
        out["speedup_vs_fp32"] = out["value"] / out["fp32_value"]
        r32 = finish(s32)
        s32.close()
        out["material_max_rel_err"] = material_err
        out["accuracy"] = {
            "steps": n_acc,
            "fields": "stored value + decoded compensation (tw_get_field_compensated), max over the 5 fields of "
                      "the relative L2 error",
            "receivers": "relative L2 error of all 256 traces over all steps",
            "reference": "FP64 run of the same scheme; 'same_material' uses the quantized coefficients of the "
                         "main run (state storage error), 'exact_material' the exact FP64 interpolation "
                         "(storage + coefficient quantization)",
            "target": 1e-4,
        }
        out["rel_l2_vs_fp64"] = {
            "storage": dict(errs(main_run, ref_q), exact_material=errs(main_run, ref_x)),
            "fp32": dict(errs(r32, ref_q), exact_material=errs(r32, ref_x)),
        }
        del main_run, r32
        # other storage variants: speed and error side by side
        out["variants"] = []
        for spec in filter(None, args.variants.split(",")):
            vp, vc, vk = (spec.split(":") + ["1"])[:3]
            vk = int(vk)
            if vk == args.coef16:
                sv = make_solver(T, W, m, T.PRECISION[vp], vc, vk, n_acc, model_from=s)
            else:  # other coefficient storage: its own quantization of the model
                sv = make_solver(T, W, m, T.PRECISION[vp], vc, vk, n_acc)
            tv, _, _ = timed_steps(torch, sv, K, Wn)
            run = finish(sv)
            # its own quantized material in FP64, where it differs from the main run's
            if vk == args.coef16:
                e_same = errs(run, ref_q)
            else:
                s64 = make_solver(T, W, m, T.FP64, "0", 0, n_acc, model_from=sv)
                e_same = errs(run, finish(s64))
                s64.close()
            sv.close()
            val = m * m * K / tv
            out["variants"].append({
                "storage": vp, "cpst": vc, "coef16": vk, "value": val, "ms_per_step": tv / K * 1e3,
                "speedup_vs_fp32": val / out["fp32_value"],
                "bytes_per_update": algorithmic_bytes(T.PRECISION[vp], vc, vk),
                "rel_l2_vs_fp64": dict(e_same, exact_material=errs(run, ref_x))})
            del run

        # configurations within the target on both counts and both materials
        def within(e):
            return max(e["fields"], e["receivers"], e["exact_material"]["fields"],
                       e["exact_material"]["receivers"]) <= out["accuracy"]["target"]
        runs = [("%s:%s:%d" % (args.precision, args.cpst, args.coef16), out["rel_l2_vs_fp64"]["storage"],
                 out["value"]), ("fp32:0:0", out["rel_l2_vs_fp64"]["fp32"], out["fp32_value"])]
        runs += [("%s:%s:%d" % (v["storage"], v["cpst"], v["coef16"]), v["rel_l2_vs_fp64"], v["value"])
                 for v in out["variants"]]
        ok = sorted((r for r in runs if within(r[1])), key=lambda r: -r[2])
        out["accuracy"]["meets_target"] = [{"config": name, "value": val, "speedup_vs_fp32": val / out["fp32_value"]}
                                           for name, _, val in ok]
    s.close()
    if not args.no_cpu:
        threads = cpu_threads()
        out["cpu_baseline"] = cpu_baseline(args.cpu_size or m)
        out["cpu_baseline"]["host_cpus"] = os.cpu_count()
        out["cpu_baseline"]["reference_acoustic"] = reference_cpu_samples(threads)
        out["cpu_baseline"]["gpu_on_reference_acoustic_case"] = gpu_reference_case(torch, T)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

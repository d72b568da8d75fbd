#!/usr/bin/env python3
"""Per-step throughput of the shell force-and-dynamics path on one MI355X.

Default workload (the metric's line): the large single shell -- unit disc,
jittered triangular lattice at spacing 0.002 (seed 0) + boundary ring,
Delaunay (~1.81M triangles, ~0.91M nodes); abar = lambda^2 I with
lambda = 1 + 0.1 g(x), g a unit-variance Gaussian random field of correlation
length 0.2 from 64 seeded Fourier modes; bbar = 0; h = 0.005, nu = 0.5,
mu = rho = 1; flat start with Gaussian z-noise (sigma 1e-4).  Synthetic data.

A step is one pass of the path over the shell: elastic forces of every
triangle, deterministic per-node gather, damping, clamp, blow-up check and
symplectic-Euler update, replayed from a HIP graph.  ``value`` is triangle
force-evaluations per second of the FP64 path (inputs resident in HBM); the
FP32-force/FP64-state mode is reported beside it.  The roofline is that of
the dominant kernel with its algorithmic bytes (see ``kern_bytes`` below).
The CPU baseline is the FP64 OpenMP oracle on the host cores, on a bounded
number of steps of the same shell.

Other workloads (one JSON line each):
  --workload batch        ragged batch of 512 discs/squares, 2k-100k triangles,
                          mixed mode, autotuned launch shapes
  --workload relax        time-to-equilibrium of the small defect disc: FP64 vs
                          mixed-then-FP64, beside the oracle on the host
  --workload relax_large  the large shell under the device-run protocol, bounded
  --workload c5           the thin-sheet stress test (warm start, then mixed), bounded
  --workload batch_relax  a 512-shell batch of 500-4000 triangles relaxed to
                          equilibrium, FP64 vs mixed-then-FP64
  --workload nn           the warm-start network: fused inference on the large
                          shell's graph, and training steps on a batch
  --workload relax_large_branch
                          the mixed relax_large run replayed to a state before
                          its blow-up, then continued in FP64 and in mixed

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_steps(ctx, states, steps: int, warmup: int, strict: bool = True) -> float:
    """Seconds for ``steps`` fused steps (after ``warmup``).  With strict=False a
    shell that blows up is halted and the others keep stepping."""
    for i, (x, v) in enumerate(states):
        ctx.set_state(i, x, v)
        ctx.resume(i)
    ctx.step(warmup, raise_on_blowup=strict)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps, raise_on_blowup=strict)
    ctx.synchronize()
    return time.perf_counter() - t0


def kernel_times(ctx, states, steps: int, strict: bool = True) -> np.ndarray:
    for i, (x, v) in enumerate(states):
        ctx.set_state(i, x, v)
        ctx.resume(i)
    ctx.enable_timing(True)
    ctx.step(steps, raise_on_blowup=strict)
    t = ctx.last_step_timing()
    ctx.enable_timing(False)
    return t


def uniform_data(shells) -> bool:
    """Every shell's thickness, shear modulus and (blend protocol) bbar_info
    are uniform over its triangles: the kernels then take them from the shell
    record and do not stream them."""
    for sh in shells:
        m = sh.mesh
        if not (np.all(m.thickness == m.thickness[0]) and np.all(m.shear_modulus == m.shear_modulus[0])):
            return False
        if sh.settings.dial_protocol != 3 and not np.all(m.bbar_info == m.bbar_info[0]):
            return False
    return True


def tri_bytes(mixed: bool, uniform: bool) -> int:
    """Algorithmic bytes of one triangle record: 6 patch ids (24 B), G and H
    (192 B FP64 / 96 B FP32), the abar dial info (24 B) and the reference
    area (8 B / 4 B); unless uniform, also bbar (24 B) and h, mu (16 B / 8 B).
    288 B FP64 and 180 B mixed in general, 248 B and 148 B with uniform h,
    mu and bbar."""
    b = 24 + (96 if mixed else 192) + 24 + (4 if mixed else 8)
    if not uniform:
        b += 24 + (8 if mixed else 16)
    return b


def kernel_bytes(T: int, N: int, mixed: bool, uniform: bool) -> list[int]:
    """Algorithmic bytes per launch of the two-launch form.  Triangle kernel:
    every triangle record once and each node position once (24 B).  Node
    kernel: x, v read and written (96 B) and the mass (8 B).  The per-slot
    contributions passed between the two launches are intermediates."""
    return [T * tri_bytes(mixed, uniform) + N * 24, N * 104]


def step_bytes(T: int, N: int, mixed: bool, uniform: bool) -> int:
    """Algorithmic bytes of one step: every triangle record once and per node
    x, v read and written plus the mass (104 B).  The fused tile kernel does
    the whole step in one launch, so this is also its per-launch figure; halo
    triangles that neighbouring tiles read again are not counted."""
    return T * tri_bytes(mixed, uniform) + N * 104


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "tile_step_grf_disc_1.8M.txt")
BATCH_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "tile_step_ragged_batch_512.txt")
# The fused step's fastest shapes on the 1.8M-triangle disc, both on 512-node
# tiles and a persistent grid of two workgroups per CU: FP64 as two chunks of
# 256 nodes (256-thread workgroups, passes of 256 triangles, so two
# workgroups fit a CU's LDS), mixed as one chunk of 512.  Tiles balanced
# over the grid's workgroups (tile_nodes=-2: 2046 tiles of 444 nodes, 4 per
# workgroup, against 1774 of 512 in 3.46 rounds) measured slower, 0.152
# against 0.144 ms in FP64 and 0.108 against 0.098 mixed: the last round's
# few workgroups run faster, and smaller tiles recompute more halo.
LARGE_FUSED = {"fp64": dict(node_block=256, tile_chunks=2, fused=1, tri_grid_per_cu=0, node_grid_per_cu=2,
                            tile_nodes=-1),
               "mixed": dict(node_block=512, tile_chunks=1, fused=1, tri_grid_per_cu=0, node_grid_per_cu=2,
                             tile_nodes=-1)}


def measured_traffic(kernel: str, profile: str = TRAFFIC_PROFILE):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel calibration), or
    None when the summary is absent."""
    try:
        with open(profile) as f:
            for line in f:
                if line.startswith("TRAFFIC_JSON "):
                    d = json.loads(line[len("TRAFFIC_JSON "):])
                    if kernel in d:
                        return dict(d[kernel], source=os.path.relpath(profile, ROOT))
    except OSError:
        pass
    return None


def device_info() -> dict:
    """The device as the HIP runtime reports it.  The roofline's peak is the
    published HBM figure: the memory clock in the device properties does not
    give HBM3E's data rate."""
    import ctypes as C

    from morphoshell_amd import native

    name = C.create_string_buffer(256)
    cus, hbm, clk, bus = C.c_int32(0), C.c_int64(0), C.c_int32(0), C.c_int32(0)
    native.check(native.lib().ms_device_info(name, 256, C.byref(cus), C.byref(hbm), C.byref(clk), C.byref(bus)))
    return {"name": name.value.decode(errors="replace"), "compute_units": cus.value, "hbm_bytes": hbm.value,
            "mem_clock_khz": clk.value, "mem_bus_bits": bus.value, "hbm_peak_GBs_published": HBM_PEAK_GBS}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(spacing: float, runs: int = 5, steps: int = 100) -> dict:
    """The FP64 oracle's step throughput on this host, as BASELINE.md plans
    it: oracle.c built here with -march=native (the reference's Makefile
    flag), run in a child process with its OpenMP threads pinned
    (OMP_PLACES=cores, OMP_PROC_BIND=spread, the reference's recommended
    launch) on every CPU this process may use; the median of `runs` runs of
    `steps` steps after warm-up, and a shorter one-thread run."""
    import subprocess
    import tempfile

    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    with tempfile.TemporaryDirectory() as tmp:
        lib = os.path.join(tmp, "liboracle_native.so")
        subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"), f"OUT={lib}", "ARCH=native"],
                       check=True, capture_output=True)
        # KMP_AFFINITY=disabled: the set-up of the shell runs the product
        # library's OpenMP (LLVM libomp), which would otherwise bind the main
        # thread to the first place; the oracle's runtime (GNU libgomp, which
        # ignores KMP_*) would then see one CPU and put its whole team there
        env = dict(os.environ, ORACLE_LIB=lib, OMP_PLACES="cores", OMP_PROC_BIND="spread",
                   OMP_NUM_THREADS=str(cpus), KMP_AFFINITY="disabled")
        r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_throughput.py"), "--spacing",
                            str(spacing), "--runs", str(runs), "--steps", str(steps)],
                           env=env, capture_output=True, text=True, timeout=900)
        if r.returncode != 0:
            raise RuntimeError("CPU baseline failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
        d = json.loads(r.stdout.strip().splitlines()[-1])
    return {
        "value": d["median"],
        "unit": "triangle force-evals/s",
        "cores": d["threads"],
        "kind": "port",
        "sample": f"median of {runs} runs of {steps} FP64 steps of the same {d['T']}-triangle shell (q=M x, dW/dq, "
                  f"F=-M^T g, damping, clamp, update) by the oracle built with -march=native, "
                  f"{d['threads']} OpenMP threads pinned (OMP_PLACES=cores, OMP_PROC_BIND=spread)",
        "runs": d["runs"],
        "one_thread": {"value": d["one_thread"], "steps": d["one_thread_steps"]},
        "cpu_model": cpu_model(),
        "logical_cpus_available": cpus,
    }


def baseline_parity(sh, x, v, n: int = 20) -> dict:
    """Part of the CPU-baseline leg: the device's forces (both modes), FP64
    energies and an n-step fused FP64 trajectory at the benchmark state,
    against the CPU baseline's own results for the same state."""
    from morphoshell_amd.shell import Context
    from oracle import oracle as orc

    parity = {}
    sysm = orc.System.from_shell(sh)
    Fo = sysm.deformation_forces(x)
    scale = np.abs(Fo).max()
    E = None
    for prec in ("fp64", "mixed"):
        with Context(sh, precision=prec) as ctx:
            ctx.set_state(0, x, v)
            ctx.calc_deformation_forces()
            parity[f"{prec}_force_rel_err"] = float(np.abs(ctx.get_forces(0) - Fo).max() / scale)
            if prec == "fp64":
                E = ctx.calc_energies(0)
    Eo = sysm.energies(x, v)
    parity["fp64_energy_rel_err"] = float(np.abs(E[:2] - Eo[:2]).max() / max(abs(Eo[:2]).max(), 1e-300))
    with Context(sh, precision="fp64") as ctx:
        ctx.set_tuning(**LARGE_FUSED["fp64"])
        ctx.set_state(0, x, v)
        ctx.step(n)
        xg = ctx.get_state(0)[0]
    xo, vo = x.copy(), v.copy()
    sysm.run_steps(xo, vo, n)
    parity[f"fp64_fused_{n}_step_position_max_abs_err"] = float(np.abs(xg - xo).max())
    return parity


# ------------------------------------------------------------------ single

def bench_single(args) -> dict:
    from morphoshell_amd import native, workloads
    from morphoshell_amd.shell import Context

    log(f"building shell, spacing {args.spacing}")
    t0 = time.time()
    sh = workloads.grf_disc(args.spacing)
    t_setup = time.time() - t0
    T, N = sh.T, sh.N
    uni = uniform_data([sh])
    log(f"T={T} N={N} mesh + set-up {t_setup:.1f}s")
    x, v = workloads.flat_noisy_state(sh)
    states = [(x, v)]
    # the product form is the fused tile step, at its fastest shape for this
    # shell; the triangle + node launch form is timed beside it for reference
    two_launch = dict(tri_block=256, node_block=256, fused=0, tri_grid_per_cu=0, node_grid_per_cu=0)
    res = {}
    for prec in args.precisions.split(","):
        with Context(sh, precision=prec) as ctx:
            for form in args.forms.split(","):
                ctx.set_tuning(**(dict(LARGE_FUSED[prec], tile_nodes=args.tile_nodes) if form == "fused"
                                  else two_launch))
                if form == "fused":
                    log(f"{prec} fused tiles of {ctx.get_tuning()['tile_nodes']} nodes")
                secs = time_steps(ctx, states, args.steps, args.warmup)
                kt = kernel_times(ctx, states, min(args.steps, 50))
                res[(prec, form)] = {"secs": secs, "kt": kt}
                log(f"{prec} {form}: {args.steps} steps {secs * 1e3:.2f} ms -> "
                    f"{T * args.steps / secs / 1e9:.3f} G tri-evals/s; kernel ms {kt[0]:.4f} + {kt[1]:.4f}")
    device = device_info()
    if not args.no_calibrate:
        cal_ms = np.zeros(1)
        native.check(native.lib().ms_calibration_copy(1 << 30, 5, native.dptr(cal_ms)))
        copy_gbs = 2 * (1 << 30) / (cal_ms[0] * 1e-3) / 1e9
        native.check(native.lib().ms_calibration_read(1 << 30, 5, native.dptr(cal_ms)))
        read_gbs = (1 << 30) / (cal_ms[0] * 1e-3) / 1e9
        device.update(copy_kernel_GBs=float(copy_gbs), read_stream_GBs=float(read_gbs))
        log(f"copy kernel 1 GiB: {copy_gbs:.0f} GB/s; read-only stream 1 GiB: {read_gbs:.0f} GB/s")
    log(f"device {device}")

    tte = None
    if not args.no_relax:
        log("time to equilibrium of the small defect disc")
        tte = c1_relaxation(args.relax_steps, warm=True, cpu_steps=2000 if not args.no_cpu else 0)

    cpu, parity = None, {}
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(args.spacing)
        log(f"cpu {cpu['value'] / 1e6:.2f} M tri-evals/s on {cpu['cores']} threads")
        if not args.no_parity:
            parity = baseline_parity(sh, x, v)
            log(f"parity {parity}")

    def summary(prec, form):
        r = res.get((prec, form))
        if r is None:
            return None
        mixed = prec == "mixed"
        ms = r["secs"] / args.steps * 1e3
        kt = r["kt"]
        d = {"value": T * args.steps / r["secs"], "ms_per_step": ms,
             "step_alg_GBs": step_bytes(T, N, mixed, uni) / (ms * 1e-3) / 1e9}
        if form == "fused":
            d["kernel_ms"] = {"tile_step": kt[0], "bookkeeping": kt[1]}
        else:
            kb = kernel_bytes(T, N, mixed, uni)
            d["kernel_ms"] = {"tri": kt[0], "node": kt[1]}
            d["tri_alg_GBs"] = kb[0] / (kt[0] * 1e-3) / 1e9
        return d

    main_prec = "fp64" if any(p == "fp64" for p, _ in res) else "mixed"
    main_form = "fused" if (main_prec, "fused") in res else "two_launch"
    r0 = res[(main_prec, main_form)]
    mixed0 = main_prec == "mixed"
    value = T * args.steps / r0["secs"]
    ms_step = r0["secs"] / args.steps * 1e3
    kt = r0["kt"]
    if main_form == "fused":
        kname = f"tile_step_kernel<{'mixed' if mixed0 else 'fp64'}> (fused forces + gather + dynamics)"
        kbytes, kms = step_bytes(T, N, mixed0, uni), kt[0]
    else:
        dom = 0 if kt[0] >= kt[1] else 1
        kname = ["tri_forces_kernel", "node_step_kernel"][dom]
        kbytes, kms = kernel_bytes(T, N, mixed0, uni)[dom], kt[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9
    traffic = measured_traffic(f"tile_step_kernel_{main_prec}") if main_form == "fused" else None
    out = {
        "metric": "triangle force-evals/s",
        "value": value,
        "unit": "triangle force-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not mixed0 else "f32 forces / f64 state",
        "data": "synthetic",
        "config": {
            "workload": f"single disc shell, spacing {args.spacing}, {T} triangles, {N} nodes, GRF lambda metric "
                        "(amp 0.1, corr 0.2, 64 modes), h=0.005, nu=0.5, FP64 forces and integration, "
                        "fused tile step",
            "num_tris": T,
            "num_nodes": N,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "traffic_detail": traffic,
            "alg_bytes_per_launch": kbytes,
            "alg_bytes_per_triangle": kbytes / T,
            "tri_record_bytes": tri_bytes(mixed0, uni),
            "uniform_h_mu_bbar": uni,
            "avg_launch_ms": kms,
        },
        "forms": {f"{p}_{f}": summary(p, f) for (p, f) in res},
        "parity": parity,
        "cpu_baseline": cpu,
        "time_to_equilibrium": tte,
        "setup_s": t_setup,
        "device": device,
    }
    return out


# ------------------------------------------------------------------- batch

def batch_traffic(tuned: dict, kernel_ms: float):
    """The fused batch kernel's HBM bytes per launch from the committed
    counter summary (same batch and tuning) over this run's kernel time."""
    if not tuned.get("fused") or tuned.get("node_block") != 512 or tuned.get("tile_chunks", 1) != 1:
        return None  # the committed counters are of 512-node tiles
    t = measured_traffic("tile_step_kernel<true, false, false, true>", BATCH_TRAFFIC_PROFILE)
    if not t:
        return None
    gbs = (t["fetch_bytes"] + t["write_bytes"]) / (kernel_ms * 1e-3) / 1e9
    return {"bytes_per_launch": t["fetch_bytes"] + t["write_bytes"], "GBs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS,
            "source": t["source"]}


def bench_batch(args) -> dict:
    from morphoshell_amd import autotune, native, workloads
    from morphoshell_amd.shell import Context

    t0 = time.time()
    shells = workloads.ragged_batch(args.shells, precision="mixed")
    t_setup = time.time() - t0
    T = sum(s.T for s in shells)
    N = sum(s.N for s in shells)
    uni = uniform_data(shells)
    log(f"batch of {len(shells)} shells: T={T} N={N}, set-up {t_setup:.1f}s")
    states = [workloads.flat_noisy_state(s, seed=i) for i, s in enumerate(shells)]
    with Context(shells, precision="mixed") as ctx:
        for i, (x, v) in enumerate(states):
            ctx.set_state(i, x, v)
        if args.no_autotune:
            # the fastest shape measured for this batch
            tuned = {"tri_block": 256, "node_block": 512, "tile_chunks": 1, "grid_per_cu": 0, "graph_steps": 16,
                     "fused": 1, "cached": None}
            ctx.set_tuning(tri_block=256, node_block=512, tri_grid_per_cu=0, node_grid_per_cu=0, graph_steps=16,
                           fused=1, tile_chunks=1)
        else:
            tuned = autotune.tune(ctx, steps=20, use_cache=False, verbose=args.verbose)
        log(f"tuning: {tuned}")
        secs = time_steps(ctx, states, args.steps, args.warmup, strict=False)
        reps = ctx.step(0, raise_on_blowup=False)
        blown = [i for i, r in enumerate(reps) if r["status"] != 0]
        T_live = T - sum(shells[i].T for i in blown)
        kt = kernel_times(ctx, states, min(args.steps, 30), strict=False)
        # relaxation of every shell within a bounded number of steps
        for i, (x, v) in enumerate(states):
            ctx.set_state(i, x, v)
            ctx.relax_setup(i, steps_per_dial_phase=1000, dial_spacing=1.0, dial_start=0.0, equil_interval=200,
                            tol_f=1e-6, tol_v=1e-5)
        t1 = time.perf_counter()
        reps = ctx.relax_run(args.batch_steps, poll_every=2000) if args.batch_steps > 0 else \
            [{"status": 0, "failed": 0, "max_force_ratio": float("nan")} for _ in shells]
        t_relax = time.perf_counter() - t1
    done = sum(1 for r in reps if r["status"] == 2)
    failed = sum(1 for r in reps if r["failed"])
    # where the bounded run leaves the shells: residual force of each one's last check
    fr = np.array([r["max_force_ratio"] for r in reps if not r["failed"]], float)
    at_tol = {f"{t:g}": int(np.sum(fr <= t)) for t in (1e-3, 1e-4, 1e-5, 1e-6)}
    kb = kernel_bytes(T, N, True, uni)
    cal = None
    if args.calibrate:
        cal_ms = np.zeros(1)
        native.check(native.lib().ms_calibration_copy(1 << 30, 5, native.dptr(cal_ms)))
        cal = {"copy_kernel_GBs": 2 * (1 << 30) / (cal_ms[0] * 1e-3) / 1e9}
        native.check(native.lib().ms_calibration_read(1 << 30, 5, native.dptr(cal_ms)))
        cal["read_stream_GBs"] = (1 << 30) / (cal_ms[0] * 1e-3) / 1e9
    return {
        "calibration": cal,
        "metric": "triangle force-evals/s (ragged batch)",
        "value": T_live * args.steps / secs,
        "shells_blown_up_in_timing": len(blown),
        "unit": "triangle force-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "ms_per_step": secs / args.steps * 1e3,
        "dtype": "fp32 forces / f64 state",
        "config": {"workload": f"ragged batch of {len(shells)} discs/squares, triangles log-uniform in [2k, 100k], "
                               "per-shell GRF lambda, bbar = kappa I, h in [0.002, 0.02]",
                   "num_shells": len(shells), "num_tris": T, "num_nodes": N},
        "autotune": tuned,
        "kernel_ms": ({"tile_step": kt[0], "bookkeeping": kt[1]} if tuned.get("fused") else
                      {"tri": kt[0], "node": kt[1]}),
        "dominant_alg_GBs": (step_bytes(T, N, True, uni) if tuned.get("fused") else kb[0]) / (kt[0] * 1e-3) / 1e9,
        "dominant_measured_traffic": batch_traffic(tuned, kt[0]),
        "step_alg_GBs": step_bytes(T, N, True, uni) / (secs / args.steps) / 1e9,
        "tri_record_bytes": tri_bytes(True, uni),
        "relax": {"steps": args.batch_steps, "seconds": t_relax, "shells_at_equilibrium": done,
                  "shells_failed": failed, "shells_per_s": done / t_relax if t_relax > 0 else 0.0,
                  "tolerances": {"force": 1e-6, "speed": 1e-5},
                  "shells_with_last_force_ratio_at_most": at_tol,
                  "last_force_ratio_median": float(np.median(fr)) if fr.size else None},
        "setup_s": t_setup,
    }


def bench_batch_relax(args) -> dict:
    """Time to equilibrium of a whole ragged batch, plain FP64 against mixed
    precision finished in FP64.  Shells as in the batch workload (discs and
    squares, per-shell random-field lambda, bbar = kappa I, h in [0.002,
    0.02]) but with triangle counts log-uniform in [batch_tmin, batch_tmax]:
    at [2k, 100k] a shell needs millions of steps.  Every shell is dialled
    over 1000 steps and checked every 200 (main.cpp:269-507 per shell); the
    mixed phase stops each shell at a residual of mixed_tol F_char (its
    forces carry an FP32 error of a few 1e-6 F_char), then the context
    switches to FP64 and every shell that did not fail is re-armed, fully
    dialled, to the real tolerances."""
    from morphoshell_amd import autotune, workloads
    from morphoshell_amd.shell import Context
    from oracle import oracle as orc

    t0 = time.time()
    shells = workloads.ragged_batch(args.shells, tmin=args.batch_tmin, tmax=args.batch_tmax, precision="mixed")
    t_setup = time.time() - t0
    T = sum(s.T for s in shells)
    N = sum(s.N for s in shells)
    states = [workloads.flat_noisy_state(s, seed=i) for i, s in enumerate(shells)]
    tol_f, tol_v, mixed_tol = 1e-6, 1e-5, 1e-5
    K, interval, chunk, cap = 1000, 200, 50_000, args.batch_relax_max_steps
    log(f"batch relax: {len(shells)} shells, T={T} N={N}, set-up {t_setup:.1f}s")

    def arm(ctx, tf, tv, start=0.0, k=K, only=None):
        for i, (x, v) in enumerate(states):
            if only is not None and i not in only:
                continue
            if start == 0.0:
                ctx.set_state(i, x, v)
            ctx.relax_setup(i, steps_per_dial_phase=k, dial_spacing=1.0, dial_start=start, equil_interval=interval,
                            tol_f=tf, tol_v=tv)

    def run(ctx, tag):
        steps, t1 = 0, time.perf_counter()
        while steps < cap:
            reps = ctx.relax_run(chunk, poll_every=10_000)
            steps += chunk
            done = sum(1 for r in reps if r["status"] == 2)
            failed = sum(1 for r in reps if r["failed"])
            log(f"{tag}: {steps} steps, {time.perf_counter() - t1:.1f}s, {done} at equilibrium, {failed} failed")
            if all(r["halted"] for r in reps):
                break
        return reps, time.perf_counter() - t1

    def energies(ctx):
        return np.array([sum(ctx.calc_energies(i)[:2]) for i in range(len(shells))])

    out = {"metric": "time-to-equilibrium (ragged batch)", "unit": "s", "higher_is_better": False,
           "config": {"workload": f"ragged batch of {len(shells)} discs/squares, triangles log-uniform in "
                                  f"[{args.batch_tmin}, {args.batch_tmax}], per-shell GRF lambda, bbar = kappa I, "
                                  f"h in [0.002, 0.02]; dialled over {K} steps, checks every {interval}; "
                                  f"force tol {tol_f}, speed tol {tol_v}",
                      "num_shells": len(shells), "num_tris": T, "num_nodes": N}}
    # plain FP64
    with Context(shells, precision="fp64") as ctx:
        arm(ctx, tol_f, tol_v)
        tuned64 = autotune.tune(ctx, steps=20, use_cache=False)
        arm(ctx, tol_f, tol_v)
        r64, t64 = run(ctx, "fp64")
        E64 = energies(ctx)
    ok64 = [i for i, r in enumerate(r64) if r["status"] == 2]
    n64 = np.array([(r["eq_step"] + 1) if r["status"] == 2 else r["steps"] for r in r64], float)
    # mixed, finished in FP64
    with Context(shells, precision="mixed") as ctx:
        arm(ctx, mixed_tol, mixed_tol)
        tunedm = autotune.tune(ctx, steps=20, use_cache=False)
        arm(ctx, mixed_tol, mixed_tol)
        rm, tm1 = run(ctx, "mixed")
        live = {i for i, r in enumerate(rm) if not r["failed"]}
        ctx.set_precision("fp64")
        arm(ctx, tol_f, tol_v, start=1.0, k=1, only=live)
        rf, tm2 = run(ctx, "mixed, fp64 finish")
        Em = energies(ctx)
    both = [i for i in ok64 if rf[i]["status"] == 2 and i in live]
    rel = np.abs(Em[both] - E64[both]) / np.abs(E64[both]) if both else np.zeros(0)
    tm = tm1 + tm2
    out["fp64"] = {"seconds": t64, "shells_at_equilibrium": len(ok64),
                   "shells_failed": sum(1 for r in r64 if r["failed"]),
                   "shells_per_s": len(ok64) / t64 if t64 > 0 else 0.0,
                   "steps_to_equilibrium": {"median": float(np.median(n64)), "max": float(n64.max())}}
    out["mixed_then_fp64"] = {"seconds": tm, "seconds_mixed": tm1, "seconds_fp64_finish": tm2,
                              "shells_at_equilibrium": sum(1 for r in rf if r["status"] == 2),
                              "shells_failed": len(shells) - len(live) + sum(1 for i in live if rf[i]["failed"]),
                              "shells_per_s": sum(1 for r in rf if r["status"] == 2) / tm if tm > 0 else 0.0,
                              "energy_rel_diff_vs_fp64": {
                                  "shells_compared": len(both),
                                  "median": float(np.median(rel)) if rel.size else None,
                                  "max": float(rel.max()) if rel.size else None,
                                  "within_1e-6": int(np.sum(rel <= 1e-6))}}
    out["speedup_mixed_vs_fp64"] = t64 / tm if tm > 0 else None
    out["tuning"] = {"fp64": tuned64, "mixed": tunedm}
    if args.batch_control:
        # control: FP64 again from the same states with a 1e-10 z-perturbation,
        # to separate a precision effect from the sensitivity of these
        # multistable shells to any rounding-level change of the trajectory
        rng = np.random.default_rng(12345)
        base = states
        states = []
        for sh, (x, v) in zip(shells, base):
            xp = x.copy()
            xp[2 * sh.N:] += 1e-10 * rng.standard_normal(sh.N)
            states.append((xp, v))
        with Context(shells, precision="fp64") as ctx:
            ctx.set_tuning(tri_block=tuned64["tri_block"], node_block=tuned64["node_block"],
                           tri_grid_per_cu=tuned64["grid_per_cu"], node_grid_per_cu=tuned64["grid_per_cu"],
                           graph_steps=tuned64["graph_steps"], fused=tuned64["fused"], tile_chunks=tuned64["tile_chunks"])
            arm(ctx, tol_f, tol_v)
            rc, tc = run(ctx, "fp64 control")
            Ec = energies(ctx)
        states = base
        bc = [i for i in ok64 if rc[i]["status"] == 2]
        relc = np.abs(Ec[bc] - E64[bc]) / np.abs(E64[bc]) if bc else np.zeros(0)
        out["control_fp64_perturbed_1e-10"] = {
            "seconds": tc, "shells_at_equilibrium": sum(1 for r in rc if r["status"] == 2),
            "energy_rel_diff_vs_fp64": {"shells_compared": len(bc),
                                        "median": float(np.median(relc)) if relc.size else None,
                                        "max": float(relc.max()) if relc.size else None,
                                        "within_1e-6": int(np.sum(relc <= 1e-6))}}
    if not args.no_cpu:
        # the oracle's per-step time of the protocol on a few shells spanning
        # the sizes, fitted as a + b T and summed over every shell's FP64
        # step count
        order = np.argsort([s.T for s in shells])
        pick = [int(order[int(q * (len(shells) - 1))]) for q in (0.0, 0.25, 0.5, 0.75, 1.0)]
        steps_cpu, ts, tt = 200, [], []
        for i in pick[:1] + pick:  # the first run warms the thread pool and is not kept
            sh = shells[i]
            sysm = orc.System.from_shell(sh)
            xx, vv = states[i][0].copy(), states[i][1].copy()
            eta = sh.scales["damping_factor"]
            tc = time.perf_counter()
            orc.run_protocol(sysm, lambda f, sh=sh: orc.dial(sh.T, 1, False, f, sh.mesh.abar_info, sh.mesh.bbar_info),
                             xx, vv, K, 1.0, 0.0, interval, eta, eta, tol_f, tol_v, steps_cpu)
            ts.append((time.perf_counter() - tc) / steps_cpu)
            tt.append(sh.T)
        ts, tt = ts[1:], tt[1:]
        b, a = np.polyfit(np.array(tt, float), np.array(ts, float), 1)
        per_shell = np.maximum(a + b * np.array([s.T for s in shells], float), min(ts))
        out["cpu_baseline"] = {"seconds_extrapolated": float(np.sum(per_shell * n64)), "cores": orc.lib().orc_num_threads(),
                               "kind": "port",
                               "sample": f"{steps_cpu} protocol steps on each of {len(pick)} shells "
                                         f"(T = {', '.join(str(t) for t in tt)}), per-step time fitted as a + b T "
                                         f"and summed over every shell's FP64 step count"}
    return out


# ------------------------------------------------------------------- relax

def bench_relax_large(args) -> dict:
    """The large smooth-random-field disc relaxed under the device-run
    protocol (dialled over 20 000 steps, checks every 1000) for a bounded
    number of steps in FP64 and in mixed mode: wall time against residual
    force, and each mode's time to the residual levels both reach.  A full
    relaxation of this shell to 1e-6 F_char is of the order of 10^7 steps.
    (Dialled over 1000 steps instead, the sheet's residual grows until the
    forces blow up, in both modes, within 25 000 steps.)"""
    from morphoshell_amd import workloads
    from morphoshell_amd.shell import Context

    sh = workloads.grf_disc(args.spacing)
    x0, v0 = workloads.flat_noisy_state(sh)
    steps, chunk, interval, K = args.relax_large_steps, 4000, 1000, 20000
    log(f"large relaxation: T={sh.T}, {steps} steps per mode")
    series, energy = {}, {}
    for prec in ("fp64", "mixed"):
        with Context(sh, precision=prec) as ctx:
            ctx.set_tuning(**LARGE_FUSED[prec])
            ctx.set_state(0, x0, v0)
            ctx.relax_setup(0, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=0.0,
                            equil_interval=interval, tol_f=1e-6, tol_v=1e-5)
            ctx.relax_run(chunk, poll_every=chunk)  # warm: graphs captured, tiles built
            ctx.set_state(0, x0, v0)
            ctx.relax_setup(0, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=0.0,
                            equil_interval=interval, tol_f=1e-6, tol_v=1e-5)
            ctx.synchronize()
            pts, t0 = [], time.perf_counter()
            while True:
                r = ctx.relax_run(chunk, poll_every=chunk)[0]
                # the residual of the last equilibrium check (none while dialling)
                f = float(r["max_force_ratio"]) if r["status"] >= 1 and r["max_force_ratio"] > 0 else float("nan")
                pts.append((time.perf_counter() - t0, int(r["steps"]), f))
                if r["halted"] or r["failed"] or r["steps"] >= steps:
                    break
            E = ctx.calc_energies(0)
        series[prec] = pts
        energy[prec] = float(E[0] + E[1])
        log(f"{prec}: {pts[-1][1]} steps in {pts[-1][0]:.2f}s, last residual {pts[-1][2]:.3e}")

    def first_time(pts, lvl):
        for t, n, f in pts:
            if f <= lvl:  # nan (no check yet) compares false
                return t, n
        return None

    levels = {}
    for lvl in (100.0, 30.0, 10.0, 3.0, 1.0, 0.3, 0.1, 0.03, 0.01, 3e-3, 1e-3, 3e-4, 1e-4):
        a, b_ = first_time(series["fp64"], lvl), first_time(series["mixed"], lvl)
        if a and b_:
            levels[f"{lvl:g}"] = {"fp64_s": a[0], "fp64_steps": a[1], "mixed_s": b_[0], "mixed_steps": b_[1],
                                  "speedup": a[0] / b_[0] if b_[0] > 0 else None}
    return {
        "metric": "time to residual force (bounded relaxation)",
        "unit": "s",
        "higher_is_better": False,
        "config": {"workload": f"single disc shell, spacing {args.spacing}, {sh.T} triangles, GRF lambda metric, "
                               f"h=0.005, dialled over {K} steps, {steps} steps per mode, fused tile step"},
        "levels_reached_by_both": levels,
        "final": {p: {"steps": series[p][-1][1], "seconds": series[p][-1][0], "force_ratio": series[p][-1][2],
                      "min_force_ratio": float(np.nanmin([f for _, _, f in series[p]] + [np.inf])),
                      "ms_per_step": 1e3 * series[p][-1][0] / max(1, series[p][-1][1]),
                      "energy": energy[p]} for p in series},
        "series": {p: [[round(t, 4), n, None if np.isnan(f) else f] for t, n, f in series[p]] for p in series},
    }


def bench_relax_large_branch(args) -> dict:
    """Why the large shell's mixed relaxation blows up.  The mixed run of
    --workload relax_large is replayed (bit for bit: the device results do not
    depend on chunking or launch shape) to the first chunk at or after
    --branch-at steps; from that state it continues once in FP64 and once in
    mixed precision (--branch-steps at most, until a blow-up), recording the
    equilibrium-check residual, the largest speed and, in the mixed branch,
    the elastic forces of the current state in both precisions (the mixed
    force error; computing them does not touch the state)."""
    from morphoshell_amd import workloads
    from morphoshell_amd.shell import Context

    sh = workloads.grf_disc(args.spacing)
    x0, v0 = workloads.flat_noisy_state(sh)
    K, interval, chunk = 20000, 1000, 4000
    fchar = sh.scales["char_force_scale"]
    out = {"workload": f"single disc shell, spacing {args.spacing}, {sh.T} triangles, mixed relaxation dialled over "
                       f"{K} steps, checks every {interval}"}
    with Context(sh, precision="mixed") as ctx:
        ctx.set_tuning(**LARGE_FUSED["mixed"])

        def setup():
            ctx.set_state(0, x0, v0)
            ctx.relax_setup(0, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=0.0, equil_interval=interval,
                            tol_f=1e-6, tol_v=1e-5)

        setup()
        ctx.relax_run(chunk, poll_every=chunk)  # warm: graphs captured, tiles built
        setup()
        ctx.set_check_log(1 << 16)
        while True:
            r = ctx.relax_run(chunk, poll_every=chunk)[0]
            if r["failed"] or r["steps"] >= args.branch_at:
                break
        recs, _ = ctx.read_check_log(0)
        res = np.array([c["max_force_ratio"] for c in recs if c["kind"] == 0 and c["step"] > K])
        out["replay"] = {"steps": int(r["steps"]), "failed": int(r["failed"]),
                         "residual_after_dial": {"median": float(np.median(res)), "max": float(res.max()),
                                                 "last": float(res[-1])} if res.size else None}
        log(f"replayed the mixed run to step {r['steps']}: {out['replay']}")
        if r["failed"]:
            return out
        snap = ctx.get_state(0)
        for name, prec in (("fp64", "fp64"), ("mixed", "mixed")):
            ctx.set_precision(prec)
            ctx.set_state(0, *snap)
            ctx.resume(0)
            start = ctx.relax_run(0, poll_every=500)[0]["steps"]
            rows = []
            while True:
                r = ctx.relax_run(500, poll_every=500)[0]
                x, v = ctx.get_state(0)
                row = {"steps": int(r["steps"] - start), "failed": int(r["failed"]),
                       "check_residual": float(r["max_force_ratio"]), "max_speed": float(np.abs(v).max())}
                if prec == "mixed" and not r["failed"]:
                    ctx.set_precision("fp64")
                    ctx.calc_deformation_forces()
                    f64 = ctx.get_forces(0)
                    ctx.set_precision("mixed")
                    ctx.calc_deformation_forces()
                    fm = ctx.get_forces(0)
                    a = float(np.abs(f64).max())
                    row.update(elastic_max_over_fchar=a / fchar,
                               mixed_err_rel=float(np.abs(fm - f64).max()) / a,
                               mixed_err_over_fchar=float(np.abs(fm - f64).max()) / fchar)
                rows.append(row)
                if r["failed"] or r["steps"] - start >= args.branch_steps:
                    break
            out[name] = {"blew_up_after": rows[-1]["steps"] if rows[-1]["failed"] else None, "rows": rows}
            log(f"{name} branch: {rows[-1]}")
    return out


def _relax_series(sh, precision, x, v, K, steps, chunk, from_ansatz=False):
    """Relax one shell under the device-run protocol (dialled over K steps,
    from its start geometry when from_ansatz) for a bounded number of steps;
    (wall seconds, steps, residual of the last check or nan) every chunk
    steps, the end state and its FP64 energies."""
    import copy
    import dataclasses

    from morphoshell_amd.shell import Context

    if from_ansatz:
        sh = copy.copy(sh)
        sh.settings = dataclasses.replace(sh.settings, dialling_from_ansatz_rather_than_ref_state=True)
    with Context(sh, precision=precision) as ctx:
        ctx.set_state(0, x, v)
        if from_ansatz:
            a, b, _ = ctx.calc_a_comps_and_b_comps_and_normals(0)
            ctx.set_zero_dial_geometry(0, a.reshape(-1), b.reshape(-1))
        ctx.set_tuning(**LARGE_FUSED[precision])
        ctx.relax_setup(0, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=0.0, equil_interval=1000,
                        tol_f=1e-6, tol_v=1e-5)
        ctx.synchronize()
        pts, t0 = [], time.perf_counter()
        while True:
            r = ctx.relax_run(chunk, poll_every=chunk)[0]
            # the residual of the last equilibrium check (none while dialling,
            # nor before the first check after it)
            f = float(r["max_force_ratio"]) if r["status"] >= 1 and r["max_force_ratio"] > 0 else float("nan")
            pts.append((time.perf_counter() - t0, int(r["steps"]), f))
            if r["halted"] or r["failed"] or r["steps"] >= steps:
                break
        xe, ve = ctx.get_state(0)
        E = ctx.calc_energies(0)
    return pts, xe, ve, E, bool(r["failed"])


def nn_counts(W: int, L: int, N: int, E: int, Fn: int, Fe: int) -> dict:
    """Algorithmic flops and bytes of one inference (ms_nn_predict) of the
    network (width W, L layers) on a graph of N nodes and E directed edges.
    Per layer, edge half: 2 E (3W W + W W) flops, reading h once (4W B per
    node), e and both node ids, writing m and e; node half: 2 N (2W W + W W)
    flops, reading h and m, writing h.  Encoders and decoder: their GEMMs."""
    edge_f, node_f = 8 * E * W * W, 6 * N * W * W
    edge_b = N * 4 * W + E * (12 * W + 8)
    node_b = N * (8 * W + 4) + E * 4 * W
    enc_f = 2 * (N * Fn + E * Fe) * W + 2 * N * (W * W + 3 * W)
    enc_b = 4 * (N * Fn + E * Fe) + 4 * (N + E) * W + 4 * N * (W + W + 3)
    return {"edge_half": {"flops": edge_f, "bytes": edge_b}, "node_half": {"flops": node_f, "bytes": node_b},
            "total_flops": L * (edge_f + node_f) + enc_f, "total_bytes": L * (edge_b + node_b) + enc_b}


def bench_nn(args) -> dict:
    """The learned warm start's network on the GPU: ms_nn_predict (fused
    half-layers) on the graph of the large single shell, and training steps
    (forward, loss, backward, Adam) on a batch of training-family graphs."""
    from morphoshell_amd import warmstart as ws
    from morphoshell_amd import workloads

    sh = workloads.grf_disc(args.spacing)
    t0 = time.perf_counter()
    g = ws.shell_graph(sh)
    b = ws.batch_graphs([g])
    t_graph = time.perf_counter() - t0
    net = ws.WarmStartNet.load(ws.PRETRAINED_DEFECT)
    W, L = net.arch
    N, E = int(b["offsets"][-1]), len(b["src"])
    counts = nn_counts(W, L, N, E, b["node_feat"].shape[1], b["edge_feat"].shape[1])
    for _ in range(2):
        net.predict(b)
    reps = max(1, args.nn_reps)
    t0 = time.perf_counter()
    for _ in range(reps):
        out = net.predict(b)
    t_pred = (time.perf_counter() - t0) / reps
    log(f"predict on N={N} E={E}: {t_pred * 1e3:.1f} ms per call (graph upload and output copy included)")
    # training steps on a batch of 8 training-family graphs
    shells = ws.training_family(8, 424_242, family="defect")
    graphs = [ws.shell_graph(s) for s in shells]
    tb = ws.batch_graphs(graphs)
    tgt = np.zeros((int(tb["offsets"][-1]), 3), np.float32)
    net.set_graph(tb)
    for _ in range(3):
        net.train_step(tgt, 1e-5)
    t0 = time.perf_counter()
    for _ in range(reps * 4):
        net.train_step(tgt, 1e-5)
    t_train = (time.perf_counter() - t0) / (reps * 4)
    net.close()
    cal = None
    if args.calibrate:
        from morphoshell_amd import native

        cal_ms = np.zeros(1)
        native.check(native.lib().ms_calibration_copy(1 << 30, 5, native.dptr(cal_ms)))
        cal = {"copy_kernel_GBs": 2 * (1 << 30) / (cal_ms[0] * 1e-3) / 1e9}
    return {
        "calibration": cal,
        "metric": "warm-start network inference (ms_nn_predict) on the large shell's graph",
        "value": t_pred * 1e3,
        "unit": "ms per prediction (graph upload and output copy included)",
        "higher_is_better": False,
        "arch": {"width": W, "layers": L},
        "graph": {"nodes": N, "edges": E, "graph_build_s": t_graph, "shell_triangles": sh.T},
        "algorithmic": counts,
        "wall_TFLOPs": counts["total_flops"] / t_pred / 1e12,
        "output_max_abs": float(np.abs(out).max()),
        "train_step_ms": t_train * 1e3,
        "train_batch": {"graphs": len(graphs), "nodes": int(tb["offsets"][-1]), "edges": len(tb["src"])},
        "kernel_profile": "profiles/nn_mlp_layer_grf_disc_1.8M.txt",
    }


def bench_c5(args) -> dict:
    """The thin-sheet stress test with the chosen parts together, bounded:
    a disc at spacing 0.001 (7.25M triangles), h = 0.002, a +1 defect metric
    (final lambda 0.9) modulated by a random field (amplitude 0.05,
    correlation length 0.1) and bbar = 1.5 I, dialled over 50 000 steps.
    Plain FP64 from the flat start with noise against the learned warm start
    (dialled from its own geometry) with mixed-precision dynamics, each run
    for the same number of steps; a relaxation to 1e-6 F_char is far beyond
    the bound.  The CPU-baseline leg checks the mixed end state against the
    FP64 oracle (forces, energies)."""
    from morphoshell_amd import warmstart as ws
    from morphoshell_amd import workloads
    from morphoshell_amd.meshing import defect_metric_abar_info, disc_mesh, grf, uniform_bbar_info
    from morphoshell_amd.shell import Context, Settings, prepare_shell

    t0 = time.time()
    m = disc_mesh(args.c5_spacing, seed=0)
    c = m.centroids()
    g = np.clip(grf(c[:, :2], 0.1, 64, seed=5), -3, 3)
    m.abar_info = defect_metric_abar_info(c, 0.9 * (1.0 + 0.05 * g), 0.5)
    m.bbar_info = uniform_bbar_info(m.num_tris, 1.5)
    m.fill_defaults(h=0.002, mu=1.0)
    sh = prepare_shell(m, Settings(def_poisson_ratio=0.5))
    t_setup = time.time() - t0
    K, steps, chunk = 50_000, args.c5_steps, 5000
    log(f"C5: T={sh.T} N={sh.N}, set-up {t_setup:.1f}s, dial over {K} steps, {steps} steps per run")
    x0, v0 = workloads.flat_noisy_state(sh)
    plain, x64, _, E64, fail64 = _relax_series(sh, "fp64", x0, v0, K, steps, chunk)
    log(f"fp64 plain: {plain[-1][1]} steps in {plain[-1][0]:.1f}s, residual {plain[-1][2]:.3e}")
    t0 = time.perf_counter()
    pred_error = None
    try:
        net = ws.WarmStartNet.load(ws.PRETRAINED_DEFECT)
        plans, used = ws.warm_start_plans(net, [sh], dial_steps=K, start="ansatz")
        net.close()  # its activations (~36 edge-sized buffers) before the dynamics
    except Exception as e:  # a failed prediction leaves the flat start
        pred_error = f"{type(e).__name__}: {e}"
        plans, used = [None], [None]
        log(f"warm start prediction failed: {pred_error}")
    t_pred = time.perf_counter() - t0
    if plans[0] is not None:
        warm, xm, vm, Em, failm = _relax_series(sh, "mixed", plans[0][0], None, K, steps, chunk, from_ansatz=True)
    else:
        warm, xm, vm, Em, failm = _relax_series(sh, "mixed", x0, v0, K, steps, chunk)
    log(f"warm start ({'used, alpha %s' % used[0] if plans[0] is not None else 'fallback to flat'}, "
        f"{t_pred:.1f}s) + mixed: {warm[-1][1]} steps in {warm[-1][0]:.1f}s, residual {warm[-1][2]:.3e}")
    uni = uniform_data([sh])
    ms64 = 1e3 * plain[-1][0] / max(1, plain[-1][1])
    msm = 1e3 * warm[-1][0] / max(1, warm[-1][1])
    out = {
        "metric": "thin-sheet stress test, bounded",
        "config": {"workload": f"disc, spacing {args.c5_spacing}, {sh.T} triangles, h=0.002, +1 defect lambda 0.9 "
                               f"x (1 + 0.05 GRF(corr 0.1)), bbar = 1.5 I, dialled over {K} steps, {steps} steps "
                               "per run"},
        "setup_s": t_setup,
        "fp64_plain": {"steps": plain[-1][1], "seconds": plain[-1][0], "ms_per_step": ms64,
                       "last_force_ratio": plain[-1][2], "failed": fail64, "energy": float(E64[0] + E64[1]),
                       "alg_GBs": step_bytes(sh.T, sh.N, False, uni) / (ms64 * 1e-3) / 1e9},
        "warm_start_mixed": {"prediction_used": plans[0] is not None, "alpha": used[0], "predict_s": t_pred,
                             "prediction_error": pred_error,
                             "steps": warm[-1][1], "seconds": warm[-1][0], "ms_per_step": msm,
                             "last_force_ratio": warm[-1][2], "failed": failm, "energy": float(Em[0] + Em[1]),
                             "alg_GBs": step_bytes(sh.T, sh.N, True, uni) / (msm * 1e-3) / 1e9},
        "energy_rel_diff_warm_mixed_vs_fp64": float((Em[0] + Em[1]) / (E64[0] + E64[1]) - 1.0),
        "series": {"fp64_plain": plain, "warm_start_mixed": warm},
    }
    out["series"] = {k: [[round(t, 3), n, None if np.isnan(f) else f] for t, n, f in v]
                     for k, v in out["series"].items()}
    if not args.no_cpu:
        # CPU-baseline leg: the FP64 oracle at the mixed run's end state
        from oracle import oracle as orc

        sysm = orc.System.from_shell(sh)
        t0 = time.perf_counter()
        Fo = sysm.deformation_forces(xm)
        t_cpu = time.perf_counter() - t0
        Eo = sysm.energies(xm, vm)
        with Context(sh, precision="mixed") as ctx:
            ctx.set_state(0, xm, vm)
            ctx.calc_deformation_forces()
            Fm = ctx.get_forces(0)
            E = ctx.calc_energies(0)
        scale = np.abs(Fo).max()
        out["oracle_check_of_mixed_end_state"] = {
            "mixed_force_rel_err": float(np.abs(Fm - Fo).max() / scale),
            "energy_rel_err": float(abs((E[0] + E[1]) - (Eo[0] + Eo[1])) / abs(Eo[0] + Eo[1])),
            "oracle_force_eval_s": t_cpu, "cores": orc.lib().orc_num_threads()}
        log(f"oracle check {out['oracle_check_of_mixed_end_state']}")
    return out


# Launch shape of the small defect disc (4510 triangles, 2255 nodes): the
# fused step on 16-node tiles (141 tiles, so most CUs get one) in 128-thread
# workgroups, one per tile.  Measured per step: FP64 11.05 us and mixed 9.84
# us, against 12.41 and 11.91 for the triangle and node launches of 64-wide
# blocks on persistent grids that it replaces.  Such shells are launch-bound.
SMALL_SHAPE = dict(tri_block=64, node_block=128, tile_chunks=1, tile_nodes=16, tri_grid_per_cu=0,
                   node_grid_per_cu=0, graph_steps=16, fused=1)


def c1_relaxation(max_steps: int, shape64: dict | None = None, shape_mixed: dict | None = None,
                  warm: bool = True, cpu_steps: int = 2000) -> dict:
    """Time to equilibrium of the small defect disc (4510 triangles, LCE
    lambda 1 -> 0.9 over 20k steps, h = 0.01) under the device-run protocol
    (main.cpp:269-507), to max |F|/F_char < 1e-6 and max |v|/c_char < 1e-5:
      * fp64: plain FP64 from the flat start with z-noise;
      * mixed_then_fp64: mixed-precision forces until the residual reaches
        1e-4 F_char (or stops falling: FP32 force noise floors it), then FP64
        to the tolerances from that state;
      * warm_mixed_then_fp64: the learned warm start (shipped parameters,
        with the energy-test fallback to the flat start), dialled from its
        own geometry, then as mixed_then_fp64.
    Each leg reports steps, seconds (prediction included), its end state's
    residual and force error against the FP64 oracle, and its elastic
    energy relative to the FP64 leg's.  The oracle runs the same protocol
    on the host for cpu_steps steps; its time is extrapolated to the FP64
    leg's step count."""
    import copy
    import dataclasses

    from morphoshell_amd import warmstart as ws
    from morphoshell_amd import workloads
    from morphoshell_amd.shell import Context
    from oracle import oracle as orc

    sh = workloads.small_defect_disc()
    x0, v0 = workloads.flat_noisy_state(sh)
    st, sc = sh.settings, sh.scales
    K = st.timesteps_per_dial_phase
    interval = max(1, int(round(st.time_between_equil_checks_prefactor * K)))
    tol_f, tol_v = st.force_to_char_force_ratio_equil_threshold, st.speed_to_char_speed_ratio_equil_threshold
    shape64 = shape64 or SMALL_SHAPE
    shape_mixed = shape_mixed or SMALL_SHAPE
    mixed_tol = 1e-4
    sysm = orc.System.from_shell(sh)
    # While waiting for equilibrium the reference keeps the targets of the last
    # dialled step, factor 1 - 1/K (do_dialling runs only while dialling,
    # main.cpp:334-353); every leg ends at those targets, and the oracle
    # checks the end states there (LCE protocol, do_dialling.cpp:104-124).
    f_end = 1.0 - 1.0 / K
    dial_end = orc.dial(sh.T, 3, False, f_end, sh.mesh.abar_info, None)
    log(f"small defect disc T={sh.T}, dt={sc['timestep']:.3e}, K={K}, check every {interval}")

    fchar = sc["char_force_scale"]

    def finish(ctx, t0, extra, targets):
        """End state of a leg (at its final equilibrium): energy, and the
        oracle's residual and force error there (in units of F_char: at an
        equilibrium max |F| itself is ~1e-6 F_char) with the leg's targets."""
        secs = time.perf_counter() - t0
        x, v = ctx.get_state(0)
        E = ctx.calc_energies(0)
        d = {"seconds": secs, "energy": float(E[0] + E[1])}
        d.update(extra)
        if extra["status"] == 2:
            ctx.set_precision("fp64")
            ctx.calc_deformation_forces()
            Fd = ctx.get_forces(0)
            sysm.set_targets(*targets)
            Fo = sysm.deformation_forces(x)
            fr, _, vr, _ = sysm.equilibrium(sysm.full_forces(x, v), v)
            d.update(oracle_force_ratio_at_end=fr, oracle_speed_ratio_at_end=vr,
                     force_err_vs_oracle_over_fchar=float(np.abs(Fd - Fo).max() / fchar))
        return d, x

    legs = {}
    # plain FP64
    with Context(sh, precision="fp64") as ctx:
        ctx.set_tuning(**shape64)
        ctx.set_state(0, x0, v0)
        ctx.relax_setup(0, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=0.0, equil_interval=interval,
                        tol_f=tol_f, tol_v=tol_v)
        t0 = time.perf_counter()
        r64 = ctx.relax_run(max_steps, poll_every=50000)[0]
        legs["fp64"], x64 = finish(ctx, t0, {
            "steps": (r64["eq_step"] + 1) if r64["status"] == 2 else r64["steps"], "status": r64["status"],
            "max_force_ratio": r64["max_force_ratio"]}, dial_end)
    log(f"fp64: {legs['fp64']}")

    def mixed_leg(shell, x, v, dial_from_ansatz, t_extra=0.0):
        with Context(shell, precision="mixed") as ctx:
            ctx.set_tuning(**shape_mixed)
            ctx.set_state(0, x, v)
            targets = dial_end
            if dial_from_ansatz:
                # the start's a, b are the targets at dial factor 0 (do_dialling.cpp:194-228)
                a, b, _ = ctx.calc_a_comps_and_b_comps_and_normals(0)
                ctx.set_zero_dial_geometry(0, a.reshape(-1), b.reshape(-1))
                targets = orc.dial(sh.T, 3, True, f_end, sh.mesh.abar_info, None, a.reshape(-1), b.reshape(-1))
            ctx.relax_setup(0, steps_per_dial_phase=K, dial_spacing=1.0, dial_start=0.0, equil_interval=interval,
                            tol_f=mixed_tol, tol_v=mixed_tol)
            t0 = time.perf_counter() - t_extra
            chunk, best = 100_000, float("inf")
            while True:
                rm = ctx.relax_run(chunk, poll_every=50000)[0]
                if rm["status"] == 2 or rm["failed"] or rm["steps"] >= max_steps:
                    break
                if rm["status"] == 1 and rm["max_force_ratio"] > 0.9 * best:
                    break
                best = min(best, rm["max_force_ratio"]) if rm["status"] == 1 else best
            # FP64 from the mixed end state: one step dials the targets to
            # f_end, then it waits for equilibrium at dial factor 1
            ctx.set_precision("fp64")
            ctx.relax_setup(0, steps_per_dial_phase=1, dial_spacing=1.0, dial_start=f_end, equil_interval=interval,
                            tol_f=tol_f, tol_v=tol_v)
            rf = ctx.relax_run(max_steps, poll_every=50000)[0]
            return finish(ctx, t0, {
                "steps_mixed": rm["steps"], "steps_fp64_finish": rf["steps"],
                "steps": rm["steps"] + ((rf["eq_step"] + 1) if rf["status"] == 2 else rf["steps"]),
                "status": rf["status"], "max_force_ratio": rf["max_force_ratio"]}, targets)

    legs["mixed_then_fp64"], xm = mixed_leg(sh, x0, v0, False)
    legs["mixed_then_fp64"]["max_position_diff_vs_fp64"] = float(np.abs(xm - x64).max())
    log(f"mixed then fp64: {legs['mixed_then_fp64']}")
    if warm:
        t0 = time.perf_counter()
        net = ws.WarmStartNet.load(ws.PRETRAINED_DEFECT)
        plans, used = ws.warm_start_plans(net, [sh], dial_steps=K, start="ansatz")
        net.close()
        t_pred = time.perf_counter() - t0
        if plans[0] is not None:
            shw = copy.copy(sh)
            shw.settings = dataclasses.replace(sh.settings, dialling_from_ansatz_rather_than_ref_state=True)
            legs["warm_mixed_then_fp64"], xw = mixed_leg(shw, plans[0][0], None, True, t_extra=t_pred)
        else:
            legs["warm_mixed_then_fp64"], xw = mixed_leg(sh, x0, v0, False, t_extra=t_pred)
        legs["warm_mixed_then_fp64"]["max_position_diff_vs_fp64"] = float(np.abs(xw - x64).max())
        legs["warm_mixed_then_fp64"].update(prediction_used=plans[0] is not None, alpha=used[0],
                                            predict_s=t_pred)
        log(f"warm start + mixed then fp64: {legs['warm_mixed_then_fp64']}")
    e64 = legs["fp64"]["energy"]
    for k in legs:
        legs[k]["energy_rel_diff_vs_fp64"] = abs(legs[k]["energy"] - e64) / abs(e64)
        legs[k]["speedup_vs_fp64"] = legs["fp64"]["seconds"] / legs[k]["seconds"]
    n64 = legs["fp64"]["steps"]
    cpu = None
    if cpu_steps > 0:  # the oracle on the host: the same protocol for cpu_steps steps
        xx, vv = x0.copy(), v0.copy()
        t0 = time.perf_counter()
        orc.run_protocol(sysm, lambda f: orc.dial(sh.T, 3, False, f, sh.mesh.abar_info, None), xx, vv, K, 1.0,
                         0.0, interval, sc["damping_factor"], sc["damping_factor"], tol_f, tol_v, cpu_steps)
        cpu_per_step = (time.perf_counter() - t0) / cpu_steps
        cpu = {"seconds_extrapolated": cpu_per_step * n64, "per_step_s": cpu_per_step,
               "cores": orc.lib().orc_num_threads(), "kind": "port",
               "sample": f"{cpu_steps} steps of the same protocol, extrapolated to the FP64 leg's {n64}"}
    return {
        "workload": f"small defect disc, {sh.T} triangles, LCE lambda 1->0.9 over {K} steps, h=0.01; "
                    f"force tol {tol_f} F_char, speed tol {tol_v} c_char, checks every {interval} steps",
        "legs": legs,
        "cpu_baseline": cpu,
        "shapes": {"fp64": shape64, "mixed": shape_mixed},
    }


def bench_relax(args) -> dict:
    from morphoshell_amd import autotune, workloads
    from morphoshell_amd.shell import Context

    sh = workloads.small_defect_disc()
    x0, v0 = workloads.flat_noisy_state(sh)
    space = dict(autotune.DEFAULT_SPACE, node_block=(64, 128), fused_node_block=(64, 128, 256),
                 fused_tile_nodes=(-2, 16, 48))
    tuned = {}
    for prec in ("fp64", "mixed"):
        with Context(sh, precision=prec) as ctx:
            ctx.set_state(0, x0, v0)
            tuned[prec] = autotune.tune(ctx, steps=400, space=space, use_cache=False)
            log(f"{prec} tuning: {tuned[prec]}")
    shapes = {p: dict({k: t[k] for k in ("tri_block", "node_block", "graph_steps", "fused", "tile_chunks",
                                         "tile_nodes") if k in t},
                      tri_grid_per_cu=t["grid_per_cu"], node_grid_per_cu=t["grid_per_cu"])
              for p, t in tuned.items()}
    res = c1_relaxation(args.relax_steps, shapes["fp64"], shapes["mixed"])
    return {"metric": "time-to-equilibrium", "unit": "s", "higher_is_better": False,
            "value": res["legs"]["fp64"]["seconds"], "config": {"workload": res["workload"]}, **res,
            "tuning": tuned}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="single",
                    choices=["single", "batch", "relax", "relax_large", "c5", "batch_relax", "nn",
                             "relax_large_branch"])
    ap.add_argument("--branch-at", type=int, default=390000, help="relax_large_branch: replay the mixed run this far")
    ap.add_argument("--branch-steps", type=int, default=12000, help="relax_large_branch: steps of each branch")
    ap.add_argument("--nn-reps", type=int, default=5, help="nn workload: timed predictions")
    ap.add_argument("--spacing", type=float, default=0.002)
    ap.add_argument("--shells", type=int, default=512)
    ap.add_argument("--relax-steps", type=int, default=3_000_000,
                    help="step bound of each time-to-equilibrium run (the small disc takes ~1M)")
    ap.add_argument("--batch-steps", type=int, default=20000,
                    help="batch workload: step bound of its bounded relaxation")
    ap.add_argument("--no-relax", action="store_true",
                    help="single workload: skip the small disc's time-to-equilibrium legs")
    ap.add_argument("--relax-large-steps", type=int, default=60000)
    ap.add_argument("--batch-tmin", type=int, default=500, help="batch_relax workload: smallest shell (triangles)")
    ap.add_argument("--batch-tmax", type=int, default=4000, help="batch_relax workload: largest shell (triangles)")
    ap.add_argument("--batch-relax-max-steps", type=int, default=4_000_000,
                    help="batch_relax workload: step bound per phase")
    ap.add_argument("--batch-control", action="store_true",
                    help="batch_relax workload: also an FP64 run from states perturbed by 1e-10")
    ap.add_argument("--c5-spacing", type=float, default=0.001)
    ap.add_argument("--c5-steps", type=int, default=60000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="batch workload: also run the copy and read kernels (the single workload always does)")
    ap.add_argument("--no-calibrate", action="store_true",
                    help="single workload: skip the copy and read kernels that measure the HBM ceiling")
    ap.add_argument("--precisions", default="fp64,mixed", help="single workload: precisions to time")
    ap.add_argument("--tile-nodes", type=int, default=-1,
                    help="single workload: nodes per fused tile (-2 balanced over the grid, -1 full chunks)")
    ap.add_argument("--forms", default="fused,two_launch", help="single workload: step forms to time")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="batch workload: the autotuner's usual pick, no search")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only (a shell, or a batch of shells, lives on one device); "
              "runs on several GPUs are not part of it.", file=sys.stderr)
        return 2
    fn = {"single": bench_single, "batch": bench_batch, "relax": bench_relax,
          "relax_large": bench_relax_large, "c5": bench_c5, "batch_relax": bench_batch_relax,
          "nn": bench_nn, "relax_large_branch": bench_relax_large_branch}[args.workload]
    print(json.dumps(fn(args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Per-step throughput of the shell force-and-dynamics path on one MI355X.

Workload (default): the large single shell -- unit disc, jittered triangular
lattice at spacing 0.002 (seed 0) + boundary ring, Delaunay (~1.81M
triangles, ~0.91M nodes); isotropic programmed metric abar = lambda^2 I with
lambda = 1 + 0.1 g(x), g a unit-variance Gaussian random field of correlation
length 0.2 from 64 seeded Fourier modes; bbar = 0; h = 0.005, nu = 0.5,
mu = rho = 1; flat start with Gaussian z-noise (sigma 1e-4).  Synthetic data.

A step is one pass of the path over the shell: elastic forces of every
triangle, deterministic per-node gather, damping, clamp, blow-up check and
symplectic-Euler update (two launches, replayed from a HIP graph).  ``value``
is triangle force-evaluations per second of the FP64 path (inputs resident in
HBM); the FP32-force/FP64-state mode is reported beside it.  The roofline of
the dominant (triangle) kernel uses its algorithmic bytes: the 288 B (FP64) /
180 B (mixed) per-triangle record plus 24 B per node of positions.  The CPU
baseline is the FP64 OpenMP oracle on the host cores, on a bounded number of
steps of the same shell.

python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def make_shell(spacing: float, seed: int = 0):
    from morphoshell_amd.meshing import disc_mesh, grf, isotropic_abar_info
    from morphoshell_amd.shell import Settings, prepare_shell

    t0 = time.time()
    m = disc_mesh(spacing, seed=seed)
    c = m.centroids()
    m.abar_info = isotropic_abar_info(1.0 + 0.1 * grf(c[:, :2], 0.2, n_modes=64, seed=seed + 1))
    m.bbar_info = np.zeros((m.num_tris, 3))
    m.fill_defaults(h=0.005, mu=1.0)
    t1 = time.time()
    sh = prepare_shell(m, Settings(def_poisson_ratio=0.5, ref_density=1.0))
    t2 = time.time()
    return sh, t1 - t0, t2 - t1


def initial_state(sh, seed: int = 0):
    rng = np.random.default_rng(seed)
    x = sh.ref_pos.copy()
    x[2 * sh.N:] += 1e-4 * rng.standard_normal(sh.N)
    return x, np.zeros_like(x)


def time_steps(ctx, x, v, steps: int, warmup: int) -> float:
    ctx.set_state(0, x, v)
    ctx.step(warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.synchronize()
    return time.perf_counter() - t0


def kernel_times(ctx, x, v, steps: int) -> np.ndarray:
    ctx.set_state(0, x, v)
    ctx.enable_timing(True)
    ctx.step(steps)
    t = ctx.last_step_timing()
    ctx.enable_timing(False)
    return t


def cpu_baseline(sh, x, v, budget_s: float) -> dict:
    from oracle import oracle as orc

    threads = orc.lib().orc_num_threads()
    sysm = orc.System.from_shell(sh)
    xx, vv = x.copy(), v.copy()
    t0 = time.perf_counter()
    sysm.run_steps(xx, vv, 1)
    t1 = time.perf_counter() - t0
    n = int(max(2, min(200, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    sysm.run_steps(xx, vv, n)
    dt = time.perf_counter() - t0
    return {
        "value": sh.T * n / dt,
        "unit": "triangle force-evals/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} FP64 steps of the same {sh.T}-triangle shell (q=M x, dW/dq, F=-M^T g, damping, "
                  f"clamp, update) with {threads} OpenMP threads",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--spacing", type=float, default=0.002)
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--calibrate", action="store_true", help="also run the copy kernel used to scale counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (each shell, or batch of shells, lives on one device); "
              "multi-GPU runs are not part of it.", file=sys.stderr)
        return 2

    from morphoshell_amd import native
    from morphoshell_amd.shell import Context

    log(f"building shell, spacing {args.spacing}")
    sh, t_mesh, t_prep = make_shell(args.spacing)
    T, N = sh.T, sh.N
    log(f"T={T} N={N} mesh {t_mesh:.1f}s set-up {t_prep:.1f}s")
    x, v = initial_state(sh)

    res = {}
    for prec in ("fp64", "mixed"):
        with Context(sh, precision=prec) as ctx:
            secs = time_steps(ctx, x, v, args.steps, args.warmup)
            kt = kernel_times(ctx, x, v, min(args.steps, 50))
            res[prec] = {"secs": secs, "kt": kt}
            log(f"{prec}: {args.steps} steps {secs * 1e3:.2f} ms -> {T * args.steps / secs / 1e9:.3f} G tri-evals/s;"
                f" kernel ms tri {kt[0]:.4f} node {kt[1]:.4f}")
            if prec == "fp64" and args.calibrate:
                cal_ms = np.zeros(1)
                native.check(native.lib().ms_calibration_copy(1 << 30, 5, native.dptr(cal_ms)))
                res["copy_GBs"] = 2 * (1 << 30) / (cal_ms[0] * 1e-3) / 1e9
                log(f"copy kernel 1 GiB: {cal_ms[0]:.3f} ms = {res['copy_GBs']:.0f} GB/s")

    parity = {}
    if not args.no_parity:
        from oracle import oracle as orc

        sysm = orc.System.from_shell(sh)
        Fo = sysm.deformation_forces(x)
        scale = np.abs(Fo).max()
        for prec in ("fp64", "mixed"):
            with Context(sh, precision=prec) as ctx:
                ctx.set_state(0, x, v)
                ctx.calc_deformation_forces()
                parity[f"{prec}_force_rel_err"] = float(np.abs(ctx.get_forces(0) - Fo).max() / scale)
                E = ctx.calc_energies(0)
        Eo = sysm.energies(x, v)
        parity["fp64_energy_rel_err"] = float(np.abs(E[:2] - Eo[:2]).max() / max(abs(Eo[:2]).max(), 1e-300))
        log(f"parity {parity}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(sh, x, v, args.cpu_budget)
        log(f"cpu {cpu['value'] / 1e6:.2f} M tri-evals/s on {cpu['cores']} threads")

    r64 = res["fp64"]
    value = T * args.steps / r64["secs"]
    ms_step = r64["secs"] / args.steps * 1e3
    kt = r64["kt"]
    dominant = 0 if kt[0] >= kt[1] else 1
    # algorithmic bytes per launch: the triangle kernel streams the 288 B
    # record of every triangle and reads each node's position once; the node
    # kernel reads x, v (48 B) and the mass (8 B) and writes x, v (48 B).
    # The per-slot contributions passed between them are intermediates.
    kern_bytes = [T * 288 + N * 24, N * 104]
    kern_names = ["tri_fp64_kernel (per-triangle forces)", "node_step_kernel (gather + dynamics)"]
    achieved = kern_bytes[dominant] / (kt[dominant] * 1e-3) / 1e9
    step_bytes = T * 288 + N * 104
    rm = res["mixed"]
    mixed_value = T * args.steps / rm["secs"]
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"single disc shell, spacing {args.spacing}, {T} triangles, {N} nodes, GRF lambda metric "
                        "(amp 0.1, corr 0.2, 64 modes), h=0.005, nu=0.5, FP64 forces and integration",
            "num_tris": T,
            "num_nodes": N,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kern_names[dominant],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "alg_bytes_per_launch": kern_bytes[dominant],
            "avg_launch_ms": kt[dominant],
        },
        "step_roofline": {
            "alg_bytes_per_step": step_bytes,
            "achieved": step_bytes / (ms_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "kernel_ms": {"tri": kt[0], "node": kt[1]},
        },
        "mixed": {
            "value": mixed_value,
            "ms_per_step": rm["secs"] / args.steps * 1e3,
            "dtype": "fp32 forces / f64 state",
            "kernel_ms": {"tri": rm["kt"][0], "node": rm["kt"][1]},
            "tri_achieved_GBs": (T * 180 + N * 24) / (rm["kt"][0] * 1e-3) / 1e9,
        },
        "parity": parity,
        "cpu_baseline": cpu,
        "setup_s": {"mesh": t_mesh, "host_precompute": t_prep},
    }
    if "copy_GBs" in res:
        out["copy_kernel_GBs"] = res["copy_GBs"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

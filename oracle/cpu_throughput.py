"""CPU baseline of the per-step path: the FP64 oracle's step throughput on
the host (test and measurement infrastructure; bench.py runs it as a child
process so that its OpenMP threads can be pinned without pinning the bench's
own threads).

    ORACLE_LIB=<oracle built with -march=native> OMP_PLACES=cores \
    OMP_PROC_BIND=spread OMP_NUM_THREADS=<n> KMP_AFFINITY=disabled \
        python oracle/cpu_throughput.py --spacing 0.002 --runs 5 --steps 100

(KMP_AFFINITY=disabled keeps the product library's OpenMP runtime, which the
shell's set-up runs first, from binding the main thread to one core: the
oracle's runtime would inherit that one-core mask.)

Builds the large smooth-field disc (the bench's default shell), starts from
the same flat state with z-noise, warms up, and times `runs` runs of `steps`
oracle steps (q = M x, dW/dq, F = -M^T g, damping, clamp, update; the
reference's dataflow, calc_deformation_forces.cpp:53-185 and
advance_dynamics.cpp:58-93) with every thread, then a shorter run on one
thread.  Prints one JSON object.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--spacing", type=float, default=0.002)
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--one-thread-steps", type=int, default=10)
    a = ap.parse_args()
    from morphoshell_amd import workloads
    from oracle import oracle as orc

    sh = workloads.grf_disc(a.spacing)
    x, v = workloads.flat_noisy_state(sh)
    L = orc.lib()
    threads = L.orc_num_threads()
    sysm = orc.System.from_shell(sh)

    def steps(n):
        if sysm.run_steps(x, v, n) != -1:
            raise RuntimeError("the oracle's step hit the blow-up test; the timing would be void")

    steps(3)  # warm-up: pages, caches, thread team
    rates = []
    for _ in range(a.runs):
        t0 = time.perf_counter()
        steps(a.steps)
        rates.append(sh.T * a.steps / (time.perf_counter() - t0))
    L.orc_set_num_threads(1)
    steps(1)
    t0 = time.perf_counter()
    steps(a.one_thread_steps)
    one = sh.T * a.one_thread_steps / (time.perf_counter() - t0)
    print(json.dumps({"T": sh.T, "threads": threads, "runs": rates, "median": float(np.median(rates)),
                      "one_thread": one, "one_thread_steps": a.one_thread_steps, "steps_per_run": a.steps,
                      "lib": os.path.basename(orc.LIB), "omp_places": os.environ.get("OMP_PLACES"),
                      "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())

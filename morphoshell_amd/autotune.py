"""Launch-shape autotuner of the per-step kernels.

Searches triangles-per-block, nodes-per-block, persistent-grid occupancy,
steps per captured HIP graph and the step form (triangle + node launches, or
the fused owner-computes tile launch) for a context, timing each candidate with the
library's own HIP events, and caches the winner by workload signature (total
triangles and nodes bucketed to powers of two, shell count bucket, precision,
device name and compute units) in a JSON file (``$MS_TUNE_CACHE`` or
``~/.cache/morphoshell_amd/tuning.json``).  Results do not depend on the
launch shape (the gather order is fixed by the CSR), so tuning never changes
numbers, only time.
"""
from __future__ import annotations

import itertools
import sys
import json
import math
import os
import time

from .native import NativeError
from .shell import Context

DEFAULT_SPACE = {
    "tri_block": (64, 128, 256),
    "node_block": (128, 256),
    "grid_per_cu": (0, 4, 8),
    "graph_steps": (16,),
    "fused_node_block": (128, 256, 512),  # block sizes of the fused step
    "fused_tile_chunks": (1, 2),      # node chunks per tile of the fused step
    # balanced tiles (cut to fill the grid, and re-cut as shells halt), or full ones
    "fused_tile_nodes": (-2, -1),
}


def _candidates(sp: dict):
    for tb, nb, g, gs in itertools.product(sp["tri_block"], sp["node_block"], sp["grid_per_cu"], sp["graph_steps"]):
        yield {"tri_block": tb, "node_block": nb, "grid_per_cu": g, "graph_steps": gs, "fused": 0}
    for nb, tc, tn, g, gs in itertools.product(sp.get("fused_node_block", ()), sp.get("fused_tile_chunks", (1,)),
                                               sp.get("fused_tile_nodes", (-1,)), sp["grid_per_cu"],
                                               sp["graph_steps"]):
        yield {"tri_block": 256, "node_block": nb, "tile_chunks": tc, "tile_nodes": tn, "grid_per_cu": g,
               "graph_steps": gs, "fused": 1}


def _apply(ctx: Context, c: dict) -> None:
    ctx.set_tuning(tri_block=c["tri_block"], node_block=c["node_block"], tri_grid_per_cu=c["grid_per_cu"],
                   node_grid_per_cu=c["grid_per_cu"], graph_steps=c["graph_steps"], fused=c.get("fused", 0),
                   tile_chunks=c.get("tile_chunks", 1), tile_nodes=c.get("tile_nodes", -1))


def _cache_path() -> str:
    p = os.environ.get("MS_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "morphoshell_amd", "tuning.json")


_DEVICE = None


def _device() -> str:
    """The device's name as the runtime reports it (its architecture name on
    parts without a marketing name), with its compute-unit count."""
    global _DEVICE
    if _DEVICE is None:
        import ctypes as C

        from . import native

        name = C.create_string_buffer(128)
        cus = C.c_int32(0)
        native.check(native.lib().ms_device_info(name, 128, C.byref(cus), None, None, None))
        dev = name.value.decode(errors="replace").split(":")[0].replace(" ", "_")
        _DEVICE = f"{dev}x{cus.value}"
    return _DEVICE


def signature(ctx: Context) -> str:
    T = sum(s.T for s in ctx.shells)
    N = sum(s.N for s in ctx.shells)
    S = len(ctx.shells)
    b = lambda v: int(round(math.log2(max(v, 1))))  # noqa: E731
    prec = "mixed" if ctx.precision == 1 else "fp64"
    return f"T2^{b(T)}-N2^{b(N)}-S2^{b(S)}-{prec}-{_device()}-v6"


def _load() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: dict) -> None:
    p = _cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p + ".tmp", "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(p + ".tmp", p)
    except OSError:
        pass


def time_config(ctx: Context, steps: int) -> float:
    """Seconds per step of the fused path under the context's current tuning."""
    ctx.step(2, raise_on_blowup=False)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.step(steps, raise_on_blowup=False)
    ctx.synchronize()
    return (time.perf_counter() - t0) / steps


def tune(ctx: Context, steps: int = 40, space: dict | None = None, use_cache: bool = True,
         verbose: bool = False) -> dict:
    """Pick and apply the fastest launch shape for ``ctx``.

    The context's state is advanced by the timing steps; callers set the
    state they want afterwards.  Returns the chosen configuration with its
    time per step.
    """
    sig = signature(ctx)
    cache = _load() if use_cache else {}
    if sig in cache:
        best = cache[sig]
        _apply(ctx, best)
        return dict(best, cached=True)
    sp = space or DEFAULT_SPACE
    results = []
    balanced_full = set()  # shapes whose balanced tiles are full ones: -1 is the same launch
    for cand in _candidates(sp):
        key = tuple(sorted((k, v) for k, v in cand.items() if k != "tile_nodes"))
        if cand.get("fused") and cand.get("tile_nodes") == -1 and key in balanced_full:
            continue
        try:
            _apply(ctx, cand)
        except NativeError as e:  # a shape the library refuses for this context (e.g. tile LDS)
            if verbose:
                print(f"[autotune] {sig} {cand}: refused ({e})", file=sys.stderr)
            continue
        if cand.get("tile_nodes") == -2 and ctx.get_tuning()["tile_nodes"] == cand["tile_chunks"] * cand["node_block"]:
            balanced_full.add(key)
        t = time_config(ctx, steps)
        results.append((t, cand))
        if verbose:
            print(f"[autotune] {sig} {cand}: {t * 1e3:.4f} ms/step", file=sys.stderr)
    t, best = min(results, key=lambda r: r[0])
    best = dict(best, ms_per_step=t * 1e3, signature=sig, candidates=len(results))
    _apply(ctx, best)
    if use_cache:
        cache[sig] = best
        _store(cache)
    return dict(best, cached=False)

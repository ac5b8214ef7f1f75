"""Autotuner of the launch options (``kah_set_option``) for one case shape.

The schedule choices of the device path do not change the arithmetic of an update (results
agree with the oracle to 1e-12 under every setting); they change how the work is mapped onto
the GPU: velocity cells per thread of the slope kernel, waves per workgroup of the flux /
relaxation kernel, wave tiles of shared velocity grids, overlap maps of adapted grids, streams
for the independent launches of a phase. The
tuner times ``kah_step`` for candidate settings by coordinate descent (one option at a time,
keeping the best value found so far) and caches the winner per case shape (dimension, grid
kind, storage type, marching, cell count and velocity-grid size buckets).
No counterpart in the reference.
"""
from __future__ import annotations

import json
import math
import time
from pathlib import Path

import numpy as np

SPACE = {
    "slope_vpt": [3, 4, 5, 6],
    "flux_waves": [0, 8, 4, 2],      # 0: the library's choice by shape
    "vel_tiles": [1, 0],
    "overlap_maps": [1, 0],
    "branches": [4, 1, 2],           # streams for the independent launches of a phase
}


def shape_key(case):
    """Cache key of a case: what the best schedule depends on."""
    s = case.config.solver
    off = np.asarray(case.topo["vs_offset"])
    nv = int(np.median(np.diff(off)))
    nc = int(case.topo["n_cells"])
    uni = int(case.topo.get("vs_uniform", 0))
    return (f"d{s.DIM}_{'uni' if uni else 'adapt'}_{getattr(case.config, 'precision', 'fp64')}_"
            f"{s.time_marching}_nc{2 ** round(math.log2(max(nc, 1)))}_"
            f"nv{2 ** round(math.log2(max(nv, 1)))}")


def relevant_space(case, space=None):
    space = dict(space or SPACE)
    uni = int(case.topo.get("vs_uniform", 0))
    if uni:
        space.pop("overlap_maps", None)     # no cross-grid reads on a shared grid
    else:
        space.pop("vel_tiles", None)        # tiles apply to shared tensor-product grids only
    return space


def time_options(case, options, steps=4, warmup=1):
    """Average ms of kah_step with the given options on a fresh context of the case."""
    from . import abi
    from .solver import kah_params
    dev = abi.HipContext(kah_params(case.config))
    try:
        for k, v in options.items():
            dev.set_option(k, v)
        dev.set_topology(case.topo)
        dev.set_state(case.w, case.prim, case.df)
        for s in range(warmup):
            dev.step(case.dt_xi, s, residual=False)
        dev.sync()
        t0 = time.perf_counter()
        for s in range(steps):
            dev.step(case.dt_xi, warmup + s, residual=False)
        dev.sync()
        return 1e3 * (time.perf_counter() - t0) / steps
    finally:
        dev.close()


def autotune(case, space=None, steps=4, warmup=1, timer=None, log=None):
    """Coordinate descent over the option space; returns (best options, best ms, trials)."""
    timer = timer or (lambda opts: time_options(case, opts, steps, warmup))
    space = relevant_space(case, space)
    best = {k: v[0] for k, v in space.items()}
    best_ms = timer(dict(best))
    trials = [(dict(best), best_ms)]
    for k, values in space.items():
        for v in values[1:]:
            cand = dict(best)
            cand[k] = v
            ms = timer(dict(cand))
            trials.append((dict(cand), ms))
            if log:
                log(f"[autotune] {cand} -> {ms:.3f} ms")
            if ms < best_ms:
                best, best_ms = cand, ms
    return best, best_ms, trials


def tuned_options(case, cache_path=None, **kw):
    """Options for the case's shape: from the cache file if present, else tuned and stored."""
    key = shape_key(case)
    path = Path(cache_path) if cache_path else None
    cache = {}
    if path and path.exists():
        cache = json.loads(path.read_text())
    if key in cache:
        return cache[key]["options"]
    best, ms, _ = autotune(case, **kw)
    if path:
        cache[key] = {"options": best, "ms_per_step": ms}
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(cache, indent=1, sort_keys=True))
    return best

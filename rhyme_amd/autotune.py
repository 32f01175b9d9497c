"""Tile-shape autotuner for the batched form.

A tile shape is (threads per workgroup, steps of 64 x-row cells per wave
along x, faces walked per thread along y/z, cap on persistent workgroups).  `tune()` times a
few steps of the real box set under each candidate (the state is restored
from the initial arrays afterwards), keeps the fastest, and caches it per
box-set signature (dimensionality, precision, limiter, and the histogram of
box sizes in powers of two) in a JSON file, so a served or repeated workload
of the same shape reuses the choice without timing again.
"""
import json
import os
import time

import numpy as np

CANDIDATES = [
    (256, 8, 8, 0), (256, 4, 8, 0), (256, 16, 8, 0), (256, 8, 16, 0), (256, 4, 4, 0),
    (128, 8, 8, 0), (128, 4, 4, 0), (64, 4, 4, 0), (256, 8, 8, 2048), (256, 4, 8, 1024),
]


def cache_path():
    p = os.environ.get("RHYME_AMD_TILE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "rhyme_amd", "tiles.json")


def signature(batch, params):
    hist = {}
    for d in batch.dims:
        key = "x".join(str(int(2 ** round(np.log2(max(n, 1))))) for n in d)
        hist[key] = hist.get(key, 0) + 1
    parts = ["nd%d" % batch.ndim, "p%d" % batch.precision, "lim%d" % params.limiter]
    parts += ["%s:%d" % (k, hist[k]) for k in sorted(hist)]
    return "|".join(parts)


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def lookup(batch, params):
    return _load().get(signature(batch, params))


def tune(batch, params, initial, steps=3, candidates=None, use_cache=True):
    """Pick the fastest tile shape for `batch` running `params`; `initial`
    is the list of host arrays to restore each box from.  On return the
    batch runs with the chosen shape and holds the initial state, whether
    the shape was timed or taken from the cache.  Returns
    (best_shape, {shape: ms_per_step})."""
    sig = signature(batch, params)
    if use_cache:
        hit = _load().get(sig)
        if hit:
            batch.set_tiles(*hit["tiles"])
            for i, c in enumerate(initial):
                batch.upload(i, c)
            return tuple(hit["tiles"]), hit.get("timings", {})
    timings = {}
    for cand in candidates or CANDIDATES:
        batch.set_tiles(*cand)
        for i, c in enumerate(initial):
            batch.upload(i, c)
        batch.run(params, 1, graph=True)  # capture + warm
        t0 = time.perf_counter()
        batch.run_async(params, steps, graph=True)
        batch.sync()
        timings[cand] = 1e3 * (time.perf_counter() - t0) / steps
    best = min(timings, key=timings.get)
    batch.set_tiles(*best)
    for i, c in enumerate(initial):
        batch.upload(i, c)
    if use_cache:
        db = _load()
        db[sig] = dict(tiles=list(best),
                       timings={",".join(map(str, k)): v for k, v in timings.items()})
        _store(db)
    return best, {",".join(map(str, k)): v for k, v in timings.items()}

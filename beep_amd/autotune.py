"""Autotuner for the enumeration launch parameters.

Sweeps, per prepared query, the minimum DFS group width (``lanes``: automatic
from the bitmap width, or 4 / 8 / 16 lanes) and the donation threshold of hub
items (|C_2| above which a level-2 subtree is donated), then, with the best of
those, the LDS bitmap build (automatic, range streaming only, neighbour search
only); each setting is timed with the library's HIP events and the fastest is
kept.  Results can be cached in a JSON file keyed by graph shape and query.
"""
import json
import os

from . import BUILD_AUTO, BUILD_SCAN, BUILD_SEARCH, EMIT_NONE

LANES = (0, 4, 8, 16)
DONATE_MIN = (16, 64, 256)
BUILD = (BUILD_AUTO, BUILD_SCAN, BUILD_SEARCH)


def tune(prepared, device=0, reps=2, lanes=LANES, donate_min=DONATE_MIN, build=BUILD):
    """Returns ({"lanes": .., "donate_min": .., "build_mode": ..}, best_ms, table) for one query."""
    def run(ln, dm, bm):
        return min(prepared.count(device=device, emit=EMIT_NONE, lanes=ln, donate=True, donate_min=dm,
                                  build_mode=bm)["t_enum_ms"] for _ in range(reps))
    table = []
    best, best_ms = None, float("inf")
    for ln in lanes:
        for dm in donate_min:
            ms = run(ln, dm, BUILD_AUTO)
            table.append({"lanes": ln, "donate_min": dm, "build_mode": BUILD_AUTO, "ms": ms})
            if ms < best_ms:
                best, best_ms = {"lanes": ln, "donate_min": dm, "build_mode": BUILD_AUTO}, ms
    for bm in build:
        if bm == BUILD_AUTO:
            continue
        ms = run(best["lanes"], best["donate_min"], bm)
        table.append({"lanes": best["lanes"], "donate_min": best["donate_min"], "build_mode": bm, "ms": ms})
        if ms < best_ms:
            best, best_ms = dict(best, build_mode=bm), ms
    return best, best_ms, table


def tune_all(prepared, device=0, cache=None, key=None, reps=2):
    """Tunes every query; with a cache path and key, reuses earlier results."""
    store = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            store = json.load(f)
    if key and key in store and len(store[key]) == len(prepared):
        return store[key]
    out = []
    for pp in prepared:
        best, ms, _ = tune(pp, device=device, reps=reps)
        out.append(dict(best, ms=ms))
    if cache and key:
        store[key] = out
        with open(cache, "w") as f:
            json.dump(store, f, indent=1)
    return out

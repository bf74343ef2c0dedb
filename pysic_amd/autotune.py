"""Autotuner for the Tersoff site kernel's launch shape.

The site kernel is compiled in several shapes (lanes per atom G and short-list
capacity); which one is fastest depends on the neighbor-count distribution
(crystal Si: ~4 bonds, amorphous: ~6-9).  best_variant() times every shape on
the loaded system with a few MD steps (HIP events around each launch on the
engine's stream) and caches the winner per (workload key, atom-count bucket)
in a JSON file next to this module.
"""
import json
import math
import os

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_cache.json")


def _bucket(n):
    return int(2 ** round(math.log2(max(n, 1))))


def _load():
    try:
        with open(CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _save(d):
    tmp = CACHE + ".tmp"
    try:
        with open(tmp, "w") as fh:
            json.dump(d, fh, indent=1, sort_keys=True)
        os.replace(tmp, CACHE)
    except OSError:
        pass


def time_variants(eng, steps=3, dt=0.0982269479):
    """Average site-kernel time (us) of every variant that runs on this
    system.  The timing steps run on the loaded neighbor lists without a
    rebuild; positions and momenta are put back afterwards, so the system
    is left as it was found (same state, same lists)."""
    out = {}
    names = eng.variants()
    pos, mom = eng.state()
    try:
        _time_each(eng, names, steps, dt, out)
    finally:
        eng.update_coordinates(pos, mom)
    return out


def _time_each(eng, names, steps, dt, out):
    for v, name in enumerate(names):
        try:
            eng.set_variant(v)
            eng.md_run(1, dt, rebuild=0)
            eng.timing(True)
            eng.md_run(steps, dt, rebuild=0)
            t = eng.timing_read()
            eng.timing(False)
        except Exception:
            eng.timing(False)
            continue
        ms, calls = t.get("tersoff_site", (0.0, 0))
        if calls:
            out[name] = ms / calls * 1e3


def best_variant(eng, key, retune=False):
    """Index of the fastest site-kernel variant for this workload (the
    system's state is unchanged by the tuning)."""
    names = eng.variants()
    ck = "%s/n%d" % (key, _bucket(eng.n))
    cache = _load()
    if not retune and ck in cache and cache[ck] in names:
        return names.index(cache[ck])
    times = time_variants(eng)
    if not times:
        return 0
    best = min(times, key=times.get)
    cache[ck] = best
    _save(cache)
    return names.index(best)

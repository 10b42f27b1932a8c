"""Autotuner of the batched solves: picks the splat tiles (tile_w x tile_h of
the pull kernels, which sets their block size, i.e. waves per block), the LDS
budget of the pulls' entry cache (which sets how many blocks, i.e. waves, a
CU holds) and the PCG iterations per graph replay, by timing a calibration
batch on the GPU it runs on.  Coordinate search: the replay chunk first, then
the tile shape, then the LDS budget, each time keeping the fastest.  Results
are cached in a JSON file keyed by the device and the workload's signature.
"""
import json
import os
import time

from . import LfrvError

TILES = [(32, 8), (64, 4), (32, 16), (16, 16), (16, 8)]
LDS_KB = [32, 48, 64]
CHUNKS = [4, 8, 16]
DEFAULT = dict(tile_w=32, tile_h=8, pull_lds_kb=48, chunk=8)


def cache_path():
    return os.environ.get("LFRV_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "lfrv", "tune.json"))


def load_cache():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(c):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            json.dump(c, f, indent=1, sort_keys=True)
    except OSError:
        pass


def tune(run_batch, signature, reps=3, use_cache=True, log=None):
    """run_batch(config) -> seconds of one batch run with that config (the
    caller builds the contexts with config's tile_w / tile_h / pull_lds_kb
    and runs a batch with its chunk).  Returns (best config, trials)."""
    cache = load_cache() if use_cache else {}
    if use_cache and signature in cache:
        return dict(cache[signature]["config"]), cache[signature]["trials"]
    trials = []

    def measure(cfg):
        ts = []
        try:
            for _ in range(reps):
                ts.append(run_batch(cfg))
        except LfrvError as e:  # a shape the library refuses for these solves
            trials.append(dict(config=dict(cfg), seconds=None, refused=str(e)))
            if log:
                log(f"autotune {cfg}: refused ({e})")
            return float("inf")
        t = min(ts)
        trials.append(dict(config=dict(cfg), seconds=t))
        if log:
            log(f"autotune {cfg}: {1e3 * t:.2f} ms")
        return t

    best = dict(DEFAULT)
    best_t = measure(best)
    for key, values in (("chunk", CHUNKS), ("tile", TILES), ("pull_lds_kb", LDS_KB)):
        for v in values:
            cfg = dict(best)
            if key == "tile":
                cfg["tile_w"], cfg["tile_h"] = v
            else:
                cfg[key] = v
            if cfg == best:
                continue
            t = measure(cfg)
            if t < best_t:
                best, best_t = cfg, t
    if use_cache:
        cache[signature] = dict(config=best, trials=trials, when=time.strftime("%Y-%m-%d %H:%M:%S"))
        save_cache(cache)
    return best, trials

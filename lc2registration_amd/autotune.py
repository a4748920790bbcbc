"""Launch-configuration autotuner with a JSON cache.

The engine's fused 3D kernel has one free launch parameter, the number of output
planes a work-group marches through (z-chunk): small chunks give more work-groups per
pose, large chunks waste fewer halo planes. ``tune(pair, mats)`` times the candidates
with HIP events inside the engine (lc2e_pair_autotune) and caches the winner per
(grid, patch radius, mode, storage, batch size) in a JSON file, so later runs with the
same shapes skip the search.
"""
from __future__ import annotations

import json
import os

DEFAULT_CACHE = os.environ.get("LC2E_TUNE_CACHE",
                               os.path.join(os.path.expanduser("~"), ".cache", "lc2e",
                                            "autotune.json"))


def _key(pair, n_poses: int) -> str:
    i = pair.info
    return (f"grid={tuple(i.virt.dims)}|r={i.patch_radius}|movi={i.high_res_is_moving}|"
            f"src={i.src_ndim}|f16={getattr(pair.moving, 'f16', 'na')}|n={n_poses}")


def load(path: str = DEFAULT_CACHE) -> dict:
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def tune(pair, mats, iters: int = 3, path: str = DEFAULT_CACHE, force: bool = False) -> int:
    """Set the pair's z-chunk for batches of len(mats) poses (cached); returns it."""
    n = len(mats)
    cache = load(path)
    key = _key(pair, n)
    if not force and key in cache:
        pair.set_tuning_n(n, int(cache[key]["zchunk"]))
        return int(cache[key]["zchunk"])
    z, ms = pair.autotune(mats, iters)
    cache[key] = {"zchunk": z, "ms_per_launch": ms}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as fh:
            json.dump(cache, fh, indent=1)
        os.replace(path + ".tmp", path)
    except OSError:
        pass
    return z

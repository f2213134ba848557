"""Launch-shape autotuner for the persistent transport kernels.

The tunable knobs are the workgroup size and the number of resident
workgroups per CU (which, for a persistent kernel, sets the number of lanes
and therefore photons per lane).  A candidate is timed with the kernel's own
HIP-event time on a sample of the workload; the winner is cached per HIP
device (its name, architecture and compute units) and workload signature in a
JSON file (XMC_AUTOTUNE_CACHE, default ~/.cache/xmcml_amd/autotune.json).
"""
from __future__ import annotations

import hashlib
import json
import os
from pathlib import Path
from typing import Callable, Iterable, List, Tuple

# (threads per workgroup, workgroups per CU; 0 = as many as fit)
CANDIDATES: List[Tuple[int, int]] = [(256, 0), (256, 4), (256, 3), (192, 0), (128, 0), (64, 0)]


def cache_path() -> Path:
    return Path(os.environ.get("XMC_AUTOTUNE_CACHE", Path.home() / ".cache" / "xmcml_amd" / "autotune.json"))


def signature(kind: str, specs) -> str:
    h = hashlib.sha256(kind.encode())
    for s in specs:
        h.update(repr((s.name, s.num_triangles, s.photons, [(l.n, l.mua, l.mus, l.g) for l in s.layers])).encode())
    return h.hexdigest()[:16]


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except Exception:
        return {}


def _store(d: dict) -> None:
    p = cache_path()
    try:
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(d, indent=1))
    except OSError:
        pass


def tune(key: str, apply: Callable[[int, int], None], time_sample: Callable[[], float],
         candidates: Iterable[Tuple[int, int]] = CANDIDATES, use_cache: bool = True) -> dict:
    """apply(block, blocks_per_cu) reconfigures the launch; time_sample() runs
    the sample and returns kernel milliseconds.  Returns the chosen shape and
    every timing."""
    cache = _load() if use_cache else {}
    if key in cache:
        b, n = cache[key]["block"], cache[key]["blocks_per_cu"]
        apply(b, n)
        return dict(cache[key], cached=True)
    timings = []
    for b, n in candidates:
        apply(b, n)
        time_sample()  # first run of a shape warms caches and the code object
        timings.append((time_sample(), b, n))
    ms, b, n = min(timings)
    apply(b, n)
    res = {"block": b, "blocks_per_cu": n, "ms": ms,
           "timings": [{"block": bb, "blocks_per_cu": nn, "ms": t} for t, bb, nn in timings]}
    if use_cache:
        cache[key] = res
        _store(cache)
    return dict(res, cached=False)


def tune_scene(scene, seed: int = 1, sample_photons: int = 200_000, use_cache: bool = True, device: int = 0) -> dict:
    from .engine import device_name
    key = f"scene:{device_name(device)}:{signature('scene', [scene.spec])}:{sample_photons}"

    def sample():
        scene.run(seed, 0, sample_photons)
        scene.sync()
        scene.collect()
        return scene.stats()["kernelMs"]

    return tune(key, scene.set_launch, sample, use_cache=use_cache)


def tune_batch(sample_batch, apply_to=None, seed: int = 1, use_cache: bool = True, device: int = 0) -> dict:
    """Tunes on `sample_batch` (the workload with scaled-down budgets) and
    applies the winner to `apply_to` as well when given."""
    from .engine import device_name
    key = f"batch:{device_name(device)}:{signature('batch', sample_batch.specs)}"

    def apply(b, n):
        sample_batch.set_launch(b, n, True)
        if apply_to is not None:
            apply_to.set_launch(b, n, True)

    def sample():
        sample_batch.run(seed)
        sample_batch.collect()
        return sample_batch.stats()["kernelMs"]

    return tune(key, apply, sample, use_cache=use_cache)

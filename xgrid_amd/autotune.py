"""Autotuner for the tile shape of temporally blocked loops.

No counterpart in the reference.  A kernel with a blockable loop can be built
with several tile shapes (core rows, region columns, work-items, iterations
per launch) and two layouts (the region in LDS, or one column per lane in
registers); which is fastest depends on the grid sizes it runs on (small
grids waste most of a large tile on halo).  ``tune_batch`` builds each
candidate, times a few batched steps on fresh copies of a sample of the real
instances (same sizes, masks and arguments; the instances themselves are not
touched) and returns the fastest, remembering the choice in a JSON file under
the kernel cache keyed by kernel, precision, device and the sample's sizes.
All candidates compute identical results; only speed differs.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

from . import _lib
from .backend import HipKernel, lower_kernel
from .emit import Emitter

CANDIDATES = [
    (48, 64, 1024, 10, "lds"),
    (48, 64, 768, 10, "lds"),
    (72, 64, 1024, 10, "lds"),
    (40, 64, 512, 10, "lds"),
    (32, 64, 512, 8, "lds"),
    (40, 64, 1024, 10, "lds"),
    (40, 64, 512, 6, "reg"),
    (40, 64, 512, 5, "reg"),
    (24, 64, 512, 4, "reg"),
]


def has_blocked_loop(op) -> bool:
    low, _ = lower_kernel(op)
    return bool(low.loops)


def _key(op, precision: str, sizes: list, candidates: list) -> str:
    from .config import get_config
    low, _ = lower_kernel(op, Emitter.DEFAULT_BLOCK_SHAPE)   # any fixed shape: the key names the kernel
    lib = _lib.load()
    import ctypes
    buf = ctypes.create_string_buffer(256)
    lib.xg_device_name(buf, 256, None, None)
    h = hashlib.sha256()
    for part in (low.device_src, low.batch_src, precision, buf.value.decode(), json.dumps(sorted(sizes)), json.dumps(candidates),
                 get_config().overstep):
        h.update(part.encode())
    return h.hexdigest()[:32]


def _cache_path() -> str:
    from .config import get_config
    root = get_config().cacheroot
    os.makedirs(root, exist_ok=True)
    return os.path.join(root, "autotune.json")


def _load() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, entry: dict) -> None:
    data = _load()
    data[key] = entry
    tmp = _cache_path() + f".tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1)
    os.replace(tmp, _cache_path())


def sample(items: list, limit: int, size=lambda x: x) -> list:
    """Up to `limit` items spread evenly over the items sorted by size."""
    if len(items) <= limit:
        return list(items)
    order = sorted(items, key=size)
    step = len(order) / limit
    return [order[int(i * step)] for i in range(limit)]


def tune_batch(op, make_instance, instances: list, precision: str, candidates: list | None = None,
               warm: int = 13, steps: int = 6, max_blocks: int = 0, use_cache: bool = True) -> dict:
    """Pick the fastest blocked-loop shape for batches of `op`.

    `make_instance(item)` returns (grids, scalars) for a fresh copy of an instance;
    `instances` are the items to time (a sample of the real batch).  The batch's
    time-level state repeats with a period of up to 6 steps and a state's graph is
    captured on its second sighting, so `warm` >= 13 steps reach steady replay and
    `steps` = 6 times one whole period.  Returns
    {"shape": ..., "timings": {shape: ms per step}, "cached": bool}."""
    from .batch import Batch
    candidates = [tuple(c) for c in (candidates or CANDIDATES)]
    sizes = [tuple(g.shape) for it in instances for g in make_instance(it)[0][:1]]
    key = _key(op, precision, sizes, [list(c) for c in candidates])
    if use_cache:
        hit = _load().get(key)
        if hit:
            return {"shape": tuple(hit["shape"]), "timings": hit["timings"], "cached": True}
    lib = _lib.load()
    timings = {}
    for shape in candidates:
        native = HipKernel(op, shape)
        if not native.lowered.loops:
            timings[",".join(map(str, shape))] = None
            native.close()
            continue
        b = Batch(op, native=native, max_blocks=max_blocks)
        for it in instances:
            grids, scalars = make_instance(it)
            b.add(*grids, *scalars)
        b.step(warm)
        _lib.check(lib.xg_synchronize())
        t0 = time.perf_counter()
        b.step(steps)
        _lib.check(lib.xg_synchronize())
        timings[",".join(map(str, shape))] = (time.perf_counter() - t0) / steps * 1e3
        b.close()
        native.close()            # its instance plans went with the batch
    valid = {k: v for k, v in timings.items() if v is not None}
    if not valid:
        return {"shape": None, "timings": timings, "cached": False}
    best = min(valid, key=valid.get)
    parts = best.split(",")
    shape = tuple(int(x) for x in parts[:4]) + (parts[4],)
    _store(key, {"shape": list(shape), "timings": timings})
    return {"shape": shape, "timings": timings, "cached": False}

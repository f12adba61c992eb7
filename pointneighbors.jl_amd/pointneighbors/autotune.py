"""Autotuner of the traversal launch and build strategy.

Knobs (pn_set_tuning):
  persist_blocks   workgroups of the persistent traversal grid (occupancy vs tail)
  chunk            LDS candidate chunk per wavefront: 0 small (more waves per CU),
                   1 large (fewer chunk passes)
  canonical_order  build strategy: 1 ranks points inside a cell by original index
                   (deterministic), 0 keeps the atomic order of the histogram pass
                   (the reference ParallelUpdate's order; only with allow_nondeterministic)

`autotune(search, step)` times `step()` (the caller's real update + traversal work) for
every combination with HIP events, keeps the fastest and records it in a JSON cache keyed
by device and workload signature, so later runs can `load_tuning` without searching.
"""

from __future__ import annotations

import itertools
import json
import os

DEFAULT_SPACE = {
    "algorithm": [0, 1],
    "persist_blocks": [2048, 8192, 32768, 131072],
    "chunk": [0, 1],
    "canonical_order": [1],
}


def signature(search, x, workload: str) -> str:
    import torch
    dev = torch.cuda.get_device_name(x.device)
    n = x.shape[1]
    bucket = 1 << max(0, int(n).bit_length() - 1)
    return f"{dev}|{search.ndims}D|{str(x.dtype).replace('torch.', '')}|n~{bucket}|{workload}"


def _time(step, reps, warmup):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        step()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune(search, step, *, space=None, reps=5, warmup=2, allow_nondeterministic=False, cache=None,
             key=None):
    """Returns (best_params, table); applies the best parameters to `search`."""
    space = dict(DEFAULT_SPACE if space is None else space)
    if allow_nondeterministic and 0 not in space.get("canonical_order", []):
        space["canonical_order"] = list(space.get("canonical_order", [1])) + [0]
    if not allow_nondeterministic:
        space["canonical_order"] = [1]
    names = list(space)
    table = []
    for combo in itertools.product(*(space[k] for k in names)):
        params = dict(zip(names, combo))
        for k, v in params.items():
            search.set_tuning(k, v)
        ms = _time(step, reps, warmup)
        table.append({"params": params, "ms": ms})
    best = min(table, key=lambda e: e["ms"])
    for k, v in best["params"].items():
        search.set_tuning(k, v)
    if cache is not None and key is not None:
        data = {}
        if os.path.exists(cache):
            with open(cache) as f:
                data = json.load(f)
        data[key] = {"params": best["params"], "ms": best["ms"], "table": table}
        tmp = cache + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, cache)
    return best["params"], table


def load_tuning(search, cache, key) -> bool:
    """Applies cached parameters for `key`; returns False when there are none."""
    if not os.path.exists(cache):
        return False
    with open(cache) as f:
        data = json.load(f)
    if key not in data:
        return False
    for k, v in data[key]["params"].items():
        search.set_tuning(k, v)
    return True

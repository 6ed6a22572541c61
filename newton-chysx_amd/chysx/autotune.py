"""Autotuner of the batched cloth substep.

Sweeps the SpMV lanes per block row (4, 8, 16 — i.e. 16, 8 or 4 block rows
per wavefront), the persistent PCG grid size (workgroups per compute unit)
and the rows per PCG tile (256, 512, 1024) on the actual batch, timing whole substeps with HIP events on the
caller's stream, and keeps the fastest.  Results are cached in a JSON file
keyed by device name and a workload signature (instance count, vertices,
off-diagonal blocks) so a later run of the same workload skips the sweep.

The sweep advances the batch's state by len(candidates) * (warmup + steps)
substeps; callers that need a pristine state tune on a copy.
"""
from __future__ import annotations

import json
import os
import time

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "chysx", "autotune.json")


def _signature(batch) -> str:
    sizes = batch.instance_sizes()
    n = sum(s[0] for s in sizes)
    nnz = sum(s[1] for s in sizes)
    return f"inst{len(sizes)}_n{n}_nnz{nnz}"


def _device_name() -> str:
    import torch

    return torch.cuda.get_device_name(0)


def tune_batch(batch, dt: float, stream=None, lanes=(4, 8, 16), blocks_per_cu=(1, 2, 3), tile_rows=(256, 512, 1024),
               warmup: int = 1, steps: int = 3, cache_path: str | None = DEFAULT_CACHE, log=None) -> dict:
    """Returns {"lanes_per_row", "grid_blocks", "tile_rows", "ms_per_substep",
    "table"} and applies the best setting to `batch`.  Lanes and grid size
    are swept at the first tile size, then the tile size at the best pair."""
    import torch

    from . import abi

    sig = _signature(batch)
    dev = _device_name()
    key = f"{dev}|{sig}"
    if cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as f:
                cached = json.load(f).get(key)
        except (OSError, ValueError):
            cached = None
        if cached:
            batch.set_lanes_per_row(int(cached["lanes_per_row"]))
            batch.set_grid_blocks(int(cached["grid_blocks"]))
            batch.set_tile_rows(int(cached.get("tile_rows", 256)))
            return dict(cached, cached=True)
    cus, cap, _ = abi.device_info()
    s = stream if stream is not None else torch.cuda.current_stream()
    sh = s.cuda_stream
    table = []

    def trial(L, g, T):
        batch.set_lanes_per_row(L)
        batch.set_grid_blocks(g)
        batch.set_tile_rows(T)
        with torch.cuda.stream(s):
            for _ in range(warmup):
                batch.step(dt, sh)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(steps):
                batch.step(dt, sh)
            e1.record(s)
            e1.synchronize()
        ms = e0.elapsed_time(e1) / steps
        table.append({"lanes_per_row": L, "grid_blocks": g, "tile_rows": T, "ms_per_substep": ms})
        if log:
            log(f"autotune lanes={L} grid={g} tile_rows={T}: {ms:.3f} ms/substep")

    for L in lanes:
        for b in blocks_per_cu:
            g = min(cap, b * cus)
            if any(r["lanes_per_row"] == L and r["grid_blocks"] == g for r in table):
                continue
            trial(L, g, tile_rows[0])
    best = min(table, key=lambda r: r["ms_per_substep"])
    for T in tile_rows[1:]:
        trial(best["lanes_per_row"], best["grid_blocks"], T)
    best = min(table, key=lambda r: r["ms_per_substep"])
    batch.set_lanes_per_row(best["lanes_per_row"])
    batch.set_grid_blocks(best["grid_blocks"])
    batch.set_tile_rows(best["tile_rows"])
    result = dict(best, table=table, device=dev, signature=sig, tuned_at=time.time())
    if cache_path:
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            data = {}
            if os.path.exists(cache_path):
                with open(cache_path) as f:
                    data = json.load(f)
            data[key] = {k: result[k] for k in ("lanes_per_row", "grid_blocks", "tile_rows", "ms_per_substep")}
            with open(cache_path, "w") as f:
                json.dump(data, f, indent=1)
        except (OSError, ValueError):
            pass
    return dict(result, cached=False)

"""Autotuner of the batched cloth substep.

Sweeps the SpMV form (LDS-staged stages or global gathers), the persistent
PCG grid size (workgroups per compute unit, up to what the form
keeps co-resident) and the rows per PCG tile (256, 512, 1024) on the actual
batch (lanes per block row only shape the CSR SpMV of caller-provided
row-major matrices; the cloth instances' sliced SpMV runs one lane per row,
so `lanes` defaults to a single value), timing whole substeps with HIP events on the
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


def tune_batch(batch, dt: float, stream=None, lanes=(4,), blocks_per_cu=(1, 2, 3, 4),
               tile_rows=(256, 512, 1024), staging=(True, False), warmup: int = 1, steps: int = 3,
               cache_path: str | None = DEFAULT_CACHE, log=None) -> dict:
    """Returns {"lds_staging", "lanes_per_row", "grid_blocks", "tile_rows",
    "ms_per_substep", "table"} and applies the best setting to `batch`.
    SpMV form, lanes and grid size (up to each form's co-residency limit)
    are swept at the first tile size, then the tile size at the best
    setting."""
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
            batch.set_lds_staging(bool(cached.get("lds_staging", True)))
            batch.set_lanes_per_row(int(cached["lanes_per_row"]))
            batch.set_grid_blocks(int(cached["grid_blocks"]))
            batch.set_tile_rows(int(cached.get("tile_rows", 256)))
            return dict(cached, cached=True)
    cus, _, _ = abi.device_info()
    s = stream if stream is not None else torch.cuda.current_stream()
    sh = s.cuda_stream
    table = []

    def apply(r):
        batch.set_lds_staging(r["lds_staging"])
        batch.set_lanes_per_row(r["lanes_per_row"])
        batch.set_grid_blocks(r["grid_blocks"])
        batch.set_tile_rows(r["tile_rows"])

    def trial(st, L, g, T):
        r = {"lds_staging": st, "lanes_per_row": L, "grid_blocks": g, "tile_rows": T}
        apply(r)
        with torch.cuda.stream(s):
            for _ in range(warmup):
                batch.step(dt, sh)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(steps):
                batch.step(dt, sh)
            e1.record(s)
            e1.synchronize()
        r["ms_per_substep"] = e0.elapsed_time(e1) / steps
        table.append(r)
        if log:
            log(f"autotune staging={int(st)} lanes={L} grid={g} tile_rows={T}: {r['ms_per_substep']:.3f} ms/substep")

    bf16 = False
    for st in staging:
        for L in lanes:
            limit = abi.pcg_grid_limit(L, bf16, st)
            for b in blocks_per_cu:
                g = min(limit, b * cus)
                if any(r["lds_staging"] == st and r["lanes_per_row"] == L and r["grid_blocks"] == g for r in table):
                    continue
                trial(st, L, g, tile_rows[0])
    best = min(table, key=lambda r: r["ms_per_substep"])
    for T in tile_rows[1:]:
        trial(best["lds_staging"], best["lanes_per_row"], best["grid_blocks"], T)
    best = min(table, key=lambda r: r["ms_per_substep"])
    apply(best)
    result = dict(best, table=table, device=dev, signature=sig, tuned_at=time.time())
    if cache_path:
        try:
            os.makedirs(os.path.dirname(cache_path), exist_ok=True)
            data = {}
            if os.path.exists(cache_path):
                with open(cache_path) as f:
                    data = json.load(f)
            data[key] = {k: result[k]
                         for k in ("lds_staging", "lanes_per_row", "grid_blocks", "tile_rows", "ms_per_substep")}
            with open(cache_path, "w") as f:
                json.dump(data, f, indent=1)
        except (OSError, ValueError):
            pass
    return dict(result, cached=False)


def tune_simulator(sim, dt: float, stream=None, staging=(True, False), tile_rows=(256, 512), forms=(0, 1),
                   grid_fractions=(1.0, 0.5), warmup: int = 1, steps: int = 3, log=None) -> dict:
    """The single-cloth path (ClothSimulator.step): sweeps the SpMV form, the
    rows per PCG tile, the solver form (one or two grid barriers per
    iteration) and the persistent grid (all tiles, or a fraction of them,
    within the form's co-residency limit), timing whole substeps with HIP
    events on the caller's stream, and applies the fastest setting to `sim`.
    A single cloth has fewer tiles than the device has CUs at most sizes, so
    the grid is bounded by the tile count rather than by occupancy.  The
    sweep advances the cloth's state by every trial's substeps.  Returns
    {"lds_staging", "tile_rows", "pcg_form", "grid_blocks", "ms_per_substep",
    "table"}."""
    import torch

    from . import abi

    s = stream if stream is not None else torch.cuda.current_stream()
    sh = s.cuda_stream
    with torch.cuda.stream(s):
        sim.step(dt, sh)  # builds the solver topology: num_particles() is 0 before it
    n = sim.num_particles()
    table = []

    def apply(r):
        sim.set_lds_staging(r["lds_staging"])
        sim.set_tile_rows(r["tile_rows"])
        sim.set_pcg_form(r["pcg_form"])
        sim.set_grid_blocks(r["grid_blocks"])

    def trial(r):
        apply(r)
        with torch.cuda.stream(s):
            for _ in range(warmup):
                sim.step(dt, sh)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(steps):
                sim.step(dt, sh)
            e1.record(s)
            e1.synchronize()
        r["ms_per_substep"] = e0.elapsed_time(e1) / steps
        table.append(r)
        if log:
            log(f"autotune (single cloth) staging={int(r['lds_staging'])} tile_rows={r['tile_rows']} "
                f"form={r['pcg_form']} grid={r['grid_blocks']}: "
                f"{r['ms_per_substep']:.3f} ms/substep")

    seen = set()
    for form in forms:
        for st in staging:
            limit = abi.pcg_grid_limit(4, False, st)
            for T in tile_rows:
                tiles = max(1, (n + T - 1) // T)
                for f in grid_fractions:
                    g = max(1, min(limit, int(round(tiles * f))))
                    if (form, st, T, g) in seen:
                        continue
                    seen.add((form, st, T, g))
                    trial({"lds_staging": st, "tile_rows": T, "pcg_form": form, "grid_blocks": g})
    best = min(table, key=lambda r: r["ms_per_substep"])
    apply(best)
    return dict(best, table=table)

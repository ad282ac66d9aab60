"""Autotuner for the tiling of the PME spreading.

The spreading kernels can cut the PME grid into 2^m tiles per axis and run
one workgroup per tile, either owning the tile's points (each written once)
or adding into an overlapping window that a second pass folds into the grid;
or they can take one atom at a time with atomic adds into the grid
(`mpid_spread_tilings` lists what a grid admits).  Which is fastest depends
on the grid and on the atoms per tile, and for the batched workload's small
systems (grids of 15-40 points, 100-3000 atoms) it is not obvious from the
LDS arithmetic alone.  `tune` times every tiling on the device for a
representative system of each grid size and writes the one whose spreading
kernels take the least device time (HIP events around them; the wall time
of a small system's evaluation is set by its launches, and in the batched
workload several systems share the device) into ``spread_tilings.json``
next to this file; `HipCalcMPIDForceKernel`
takes its tiling from that table when the grid is in it, and leaves the
choice to the library otherwise.

Run on the GPU (one MI355X)::

    python -m mpidopenmmplugin_amd.autotune            # writes spread_tilings.json

The reference has no counterpart: its spreading is one loop over the atoms
(MPIDReferenceForce.cpp:3269-3327, 3532-3573).
"""
from __future__ import annotations

import json
import math
import os
import time

from . import _capi

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spread_tilings.json")
_MODES = {0: "untiled", 1: "owner", 2: "windows"}
_table = None


def grid_key(grid):
    return ",".join(str(int(g)) for g in grid)


def load_table(path=TABLE_PATH):
    """The tuned table ({grid key: entry}), empty if there is none."""
    global _table
    if path != TABLE_PATH:
        with open(path) as fh:
            return json.load(fh).get("entries", {})
    if _table is None:
        _table = {}
        if os.path.exists(path):
            with open(path) as fh:
                _table = json.load(fh).get("entries", {})
    return _table


def lookup(grid):
    """The tuned (spread_level, spread_mode) for a PME grid, or None (the
    library's own choice) when the grid is not in the table or its entry is
    not among the tilings the library offers for it."""
    e = load_table().get(grid_key(grid))
    if e is None:
        return None
    pick = (int(e["level"]), int(e["mode"]))
    try:
        return pick if pick in _capi.spread_tilings(grid) else None
    except (OSError, ImportError, ValueError):
        return None


def _systems_by_grid(grids):
    """One representative system per grid size: the batched workload's
    generator at the largest water count (32-1024) that gives the grid, the
    4096- and 32768-water boxes for 64 and 128."""
    from .waterbox import WATERS_PER_NM3, make_small_waterbox, make_waterbox, smooth_235

    def grid_of(nw):
        L = (nw / WATERS_PER_NM3) ** (1.0 / 3.0)
        return max(6, smooth_235(math.ceil(L / 0.08)))

    out = {}
    for g in grids:
        if g == 64:
            out[g] = make_waterbox("water4096")
        elif g == 128:
            out[g] = make_waterbox("water32768")
        else:
            nws = [nw for nw in range(32, 1025) if grid_of(nw) == g]
            if nws:
                out[g] = make_small_waterbox(max(nws), seed=1000 + g)
    return out


def tune(grids=(15, 16, 18, 20, 24, 25, 27, 30, 32, 36, 40, 64, 128), precision="mixed", reps=30, warmup=3,
         path=TABLE_PATH, log=print):
    """Time every tiling of each grid on cuda:0 and write the fastest to path."""
    import torch

    from .kernel import HipCalcMPIDForceKernel

    dev = torch.device("cuda:0")
    entries = {}
    for g, (force, pos, box) in _systems_by_grid(grids).items():
        grid = (g, g, g)
        d_pos = torch.tensor(pos, dtype=torch.float64, device=dev).contiguous()
        wall, gpu, energies = {}, {}, {}
        for level, mode in _capi.spread_tilings(grid):
            k = HipCalcMPIDForceKernel(precision, spread_tiling=(level, mode))
            k.initialize(force.getNumMultipoles(), force, default_box=box)
            for _ in range(warmup):
                k.execute_device(d_pos.data_ptr(), box)
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(reps):
                e = k.execute_device(d_pos.data_ptr(), box)
            torch.cuda.synchronize()
            name = f"{level},{mode}"
            wall[name] = 1000.0 * (time.perf_counter() - t) / reps
            energies[name] = e
            # the criterion: device time of the spreading kernels (grid
            # clearing, spreading and folding, every spread of an
            # evaluation), from HIP events around them
            k.setProfileFamilies(["spread"])
            k.profile(reset=True)
            for _ in range(reps):
                k.execute_device(d_pos.data_ptr(), box)
            torch.cuda.synchronize()
            gpu[name] = k.profile()["spread"][0] / reps
            del k
        best = min(gpu, key=gpu.get)
        lv, md = (int(x) for x in best.split(","))
        e0 = energies[best]
        spread = max(abs(v - e0) for v in energies.values()) / max(abs(e0), 1e-30)
        entries[grid_key(grid)] = {"level": lv, "mode": md, "mode_name": _MODES[md], "atoms": int(len(pos)),
                                   "spread_ms_per_evaluation": gpu, "ms_per_evaluation": wall,
                                   "energy_rel_spread": spread}
        log(f"[autotune] grid {g}^3, {len(pos)} atoms: spreading " +
            ", ".join(f"({k}) {v:.3f} ms" for k, v in gpu.items()) + "; evaluation " +
            ", ".join(f"({k}) {v:.3f} ms" for k, v in wall.items()) +
            f" -> ({best}), energies agree to {spread:.1e}")
    table = {"precision": precision, "device": torch.cuda.get_device_name(0), "reps": reps,
             "command": "python -m mpidopenmmplugin_amd.autotune", "entries": entries}
    with open(path, "w") as fh:
        json.dump(table, fh, indent=1, sort_keys=True)
    global _table
    _table = None
    return table


if __name__ == "__main__":
    tune(log=lambda s: print(s, flush=True))

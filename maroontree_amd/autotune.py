"""Autotuner of the transform chain's per-size launch schedules.

The chain (mt_tx_chain) runs one persistent kernel per transform size over that size's
blocks.  Two shape parameters of each kernel are free, and the results never depend on
them (every line loop strides over the tile's blocks, and each block is finished by the
workgroup that holds it):

- the tile: how many blocks a 256-thread workgroup stages in LDS (tile_shift -1 / 0 / 1:
  half, default, double the default max(256 / min(W, H), 4096 / (W * H))), which trades
  LDS per workgroup (occupancy) against lines per pass;
- the persistent grid: workgroups per CU (for 64x64, one-wave workgroups per CU).

`tune()` times every candidate on a batch of each size alone (the size's blocks of the
microbench workload: its type mix, Laplace residuals, bit depths and quantizers) with HIP
events on the launching stream, keeps the fastest, installs it with mt_tx_schedule_set
and returns the table; `save()` / `apply()` keep it as JSON next to the library.
"""
import json
import os
import time

import numpy as np

from . import _lib as L

PKG = os.path.dirname(os.path.abspath(__file__))
SCHEDULE_FILE = os.path.join(PKG, "tx_schedule.json")

TILE_CANDIDATES = [(ts, wg) for ts in (-1, 0, 1) for wg in (2, 4, 8, 16)]
WAVE_CANDIDATES = [(0, wg) for wg in (8, 16, 32, 64)]


def candidates(tx_size):
    """The (tile_shift, wg_per_cu) schedules tried for one transform size."""
    return WAVE_CANDIDATES if tx_size == 4 else TILE_CANDIDATES


def _time_chain(bt, pt, rt, stream, reps):
    import torch
    L.tx_chain(bt, pt, rt, stream=stream)  # warm-up (workspace growth, code load)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    for r in range(reps):
        ev[2 * r].record(stream)
        L.tx_chain(bt, pt, rt, stream=stream)
        ev[2 * r + 1].record(stream)
    stream.synchronize()
    return float(np.median([ev[2 * r].elapsed_time(ev[2 * r + 1]) for r in range(reps)]))


def tune(device=0, n=1 << 20, seed=0x6D61726F6F6E, reps=5, log=None, workload=None):
    """Measures every candidate schedule of the 13 chain sizes on `device` and installs the
    fastest; returns {"sizes": {name: {...}}, ...}.  The batch is the microbench workload
    of n blocks, split by size, or `workload` = (host blocks, device int16 pool) given."""
    import torch
    L.init(device)
    dev = torch.device("cuda", device)
    stream = torch.cuda.Stream(device=dev)
    if workload is None:
        blocks, pool = L.gen_tx_blocks(seed, n)
        pt = torch.from_numpy(pool).to(dev)
    else:
        blocks, pt = workload
        n = len(blocks)
    t0 = time.time()
    out = {"device": torch.cuda.get_device_name(dev), "n_workload": n, "seed": seed, "reps": reps,
           "sizes": {}}
    with torch.cuda.stream(stream):
        for sz in L.TX_SIZES_REF:
            sel = blocks[blocks["tx_size"] == sz]
            if len(sel) == 0:
                continue
            bt = torch.from_numpy(sel.view(np.uint8).copy()).to(dev)
            rt = torch.empty(len(sel) * L.RES_DTYPE.itemsize, dtype=torch.uint8, device=dev)
            default = L.tx_schedule_get(sz, device)
            times = {}
            for ts, wg in candidates(sz):
                L.tx_schedule_set(sz, ts, wg, device)
                times[(ts, wg)] = _time_chain(bt, pt, rt, stream, reps)
            best = min(times, key=times.get)
            L.tx_schedule_set(sz, best[0], best[1], device)
            rec = {"tx_size": sz, "blocks": int(len(sel)), "tile_shift": best[0], "wg_per_cu": best[1],
                   "ms": round(times[best], 4), "default": list(default),
                   "default_ms": round(times.get(tuple(default), float("nan")), 4),
                   "tried": {f"{a},{b}": round(v, 4) for (a, b), v in sorted(times.items())}}
            out["sizes"][L.TX_NAMES[sz]] = rec
            if log:
                log(f"autotune {L.TX_NAMES[sz]}: {len(sel)} blocks, best tile_shift {best[0]} "
                    f"wg/CU {best[1]} {times[best]:.3f} ms (default {rec['default_ms']:.3f} ms)")
    out["seconds"] = round(time.time() - t0, 2)
    return out


def save(table, path=SCHEDULE_FILE):
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def apply(table=None, device=0, path=SCHEDULE_FILE):
    """Installs a saved (or given) schedule table on `device`; False when there is none."""
    if table is None:
        if not os.path.exists(path):
            return False
        with open(path) as f:
            table = json.load(f)
    for rec in table["sizes"].values():
        L.tx_schedule_set(rec["tx_size"], rec["tile_shift"], rec["wg_per_cu"], device)
    return True

"""Autotuner for the tiled kernel's shape knobs.

The knobs (points per lane with its default tile, LDS list capacity, Widom
bin occupancy, region edge, and one tile shape per grid size class) decide
how a batch maps onto workgroups; none of them changes the physics, and each
framework gets the same knobs alone, in a batch or in the server.  The
search is coordinate descent over measured batch time: one knob at a time,
every candidate of that knob timed with the others held, the best kept if it
beats the incumbent by more than the noise margin, until a full sweep
changes nothing.  Timings are memoised, so revisiting a configuration costs
nothing.

Size classes are by grid point spacing (the geometric mean of the three
lattice-direction spacings): class 0 below 0.15 A, class 1 below 0.35 A,
class 2 the rest, matching the 0.1 / 0.2 / 0.5 A grids of the batching mix.

The result is a JSON file holding the knobs and the search log; `load`
turns it back into a Tuning for a context.
"""
import json
import statistics
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from . import api

POW2 = (1, 2, 4, 8, 16, 32, 64)


def tiles_for(ppl: int) -> List[Tuple[int, int, int]]:
    """Tile shapes of 256 * ppl points with power-of-two extents in [2, 32]
    whose y-z face divides 256 (a lane's points then differ only in i)."""
    n = 256 * ppl
    out = []
    for a in POW2:
        for b in POW2:
            if n % (a * b):
                continue
            c = n // (a * b)
            if all(2 <= x <= 32 for x in (a, b, c)) and 256 % (b * c) == 0:
                out.append((a, b, c))
    return out


PPL_TILES = [(1, (8, 8, 4)), (2, (8, 8, 8)), (4, (16, 8, 8))]
LDS_RECORDS = [768, 1024, 1280, 1536]
WIDOM_PER_BIN = [240, 360, 480, 720, 960]
REGION_LEN = [0.0, 2.5, 4.0, 6.0]


def _key(cfg: dict) -> str:
    return json.dumps(cfg, sort_keys=True)


def knobs(classes: Sequence[int]) -> List[Tuple[str, Callable[[dict], list]]]:
    """(name, candidates(cfg)) in search order.  Changing points per lane
    resets the class tiles (they must hold 256 * ppl points)."""
    def ppl_tile(cfg):
        out = []
        for p, t in PPL_TILES:
            c = dict(cfg, points_per_lane=p, tile=list(t),
                     class_tile=[[0, 0, 0], [0, 0, 0], [0, 0, 0]])
            out.append(c)
        return out

    def simple(name, values):
        return lambda cfg: [dict(cfg, **{name: v}) for v in values]

    def class_tile(c):
        def cand(cfg):
            out = [_with_class(cfg, c, [0, 0, 0])]
            for t in tiles_for(cfg["points_per_lane"]):
                if list(t) != list(cfg["tile"]):
                    out.append(_with_class(cfg, c, list(t)))
            return out
        return cand

    ks = [("ppl_tile", ppl_tile), ("lds_records", simple("lds_records", LDS_RECORDS)),
          ("widom_per_bin", simple("widom_per_bin", WIDOM_PER_BIN)),
          ("region_len", simple("region_len", REGION_LEN))]
    for c in sorted(set(classes)):
        ks.append((f"class_tile[{c}]", class_tile(c)))
    return ks


def _with_class(cfg: dict, c: int, tile: list) -> dict:
    ct = [list(r) for r in cfg["class_tile"]]
    ct[c] = list(tile)
    return dict(cfg, class_tile=ct)


def coordinate_search(measure: Callable[[dict], float], start: dict,
                      knob_list: List[Tuple[str, Callable[[dict], list]]],
                      rounds: int = 2, margin: float = 0.01,
                      log: Optional[Callable[[str], None]] = None):
    """Minimise measure(cfg) one knob at a time.  A candidate replaces the
    incumbent only if it is faster by more than `margin` (relative).
    Returns (best cfg, its time, [(cfg, time), ...] in measurement order)."""
    cache: Dict[str, float] = {}
    history = []

    def timed(cfg):
        k = _key(cfg)
        if k not in cache:
            cache[k] = float(measure(cfg))
            history.append((cfg, cache[k]))
            if log:
                log(f"{cache[k] * 1e3:9.2f} ms  {k}")
        return cache[k]

    best = dict(start)
    best_t = timed(best)
    for _ in range(rounds):
        changed = False
        for name, cands in knob_list:
            for cfg in cands(best):
                t = timed(cfg)
                if t < best_t * (1.0 - margin):
                    best, best_t, changed = cfg, t, True
        if not changed:
            break
    return best, best_t, history


def batch_timer(ctx, tasks: List[dict], repeats: int = 3) -> Callable[[dict], float]:
    """measure(cfg): median wall time of one batch step of `tasks` on ctx
    with knobs cfg (a fresh batch per configuration, one untimed run)."""
    base = api.default_tuning()

    def measure(cfg):
        ctx.tuning = api.Tuning.from_dict(cfg, base=base)
        b = api.Batch(ctx, tasks, use_graph=False)
        try:
            b.run()
            b.wait()
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                b.run()
                b.wait()
                ts.append(time.perf_counter() - t0)
        finally:
            b.close()
        return statistics.median(ts)
    return measure


def size_classes(tuning: "api.Tuning", grids: Sequence[Tuple[object, Sequence[int]]]):
    """Size class of each (t_matrix, n) under `tuning`."""
    return [api.grid_tile_for(tuning, t, n)[1] for t, n in grids]


def autotune(ctx, tasks: List[dict], cells: Sequence[object], rounds: int = 2,
             repeats: int = 3, margin: float = 0.01, log=None) -> dict:
    """Tune ctx's knobs on a batch of `tasks` (cells[i] = cell matrix of
    tasks[i]'s framework).  Leaves the best knobs set on ctx and returns
    {"tuning": ..., "seconds": ..., "history": [...]}."""
    start = ctx.tuning.to_dict()
    classes = size_classes(ctx.tuning, [(c, t["grid_n"]) for c, t in zip(cells, tasks)
                                        if t.get("grid_n")])
    t0 = time.perf_counter()
    best, best_t, hist = coordinate_search(batch_timer(ctx, tasks, repeats), start,
                                           knobs(classes), rounds, margin, log)
    ctx.tuning = api.Tuning.from_dict(best, base=api.default_tuning())
    return {"tuning": best, "step_seconds": best_t, "start": start,
            "start_step_seconds": hist[0][1], "search_seconds": time.perf_counter() - t0,
            "measurements": len(hist),
            "history": [{"tuning": c, "step_seconds": t} for c, t in hist]}


def save(path: str, result: dict, meta: Optional[dict] = None) -> None:
    out = dict(result)
    if meta:
        out["meta"] = meta
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")


def load(path: str) -> "api.Tuning":
    """Tuning from a file written by save (unknown knobs are an error)."""
    with open(path) as fh:
        d = json.load(fh)
    return api.Tuning.from_dict(d["tuning"], base=api.default_tuning())

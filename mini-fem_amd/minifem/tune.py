"""Autotuner for the tiled gather schedule.

For a given mesh it builds the plan for each candidate (LDS budget per tile, threads per
workgroup, row cap per tile), times the fused kernel with HIP events on the device, and
returns the fastest. Results are cached in tuned.json next to this file, keyed by operator,
precision and mesh size, so later runs reuse them without re-timing.
"""
import json
import os
import time

from . import gpu

TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")

# small budgets give small meshes enough tiles to fill 256 CUs; large ones amortize the
# per-tile chain on large meshes
DEFAULT_CANDIDATES = [
    dict(tile_lds_bytes=lds, block_threads=bt, max_tile_rows=254)
    for lds, bt in ((16 * 1024, 512), (24 * 1024, 512), (32 * 1024, 512), (48 * 1024, 512),
                    (64 * 1024, 512), (78 * 1024, 512), (100 * 1024, 1024),
                    (120 * 1024, 1024), (150 * 1024, 1024), (155000, 1024))
]


def _key(mesh, precision):
    return f"op{mesh.operatorID}_p{precision}_n{mesh.nbNodes}_e{mesh.nbElem}"


def load_tuned(mesh, precision):
    if not os.path.exists(TUNED_PATH):
        return None
    with open(TUNED_PATH) as f:
        return json.load(f).get(_key(mesh, precision))


def time_plan(plan, dm, precision, iters, stream):
    for _ in range(2):
        plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision, stream=stream)
    stream.sync()
    e0, e1 = gpu.Event(), gpu.Event()
    e0.record(stream)
    for _ in range(iters):
        plan.assemble(dm.coord, dm.coef, dm.values, dm.prec, precision=precision, stream=stream)
    e1.record(stream)
    return e0.elapsed_ms(e1) / iters * 1e3


def tune(mesh, dm, precision=0, candidates=None, iters=10, save=True, log=None):
    """Times every candidate; returns (best_params, [(params, us, stats), ...]). For a
    mixed precision mode the candidates' tiles are sized by the float stage unless a
    candidate names its stage_precision."""
    cands = candidates or DEFAULT_CANDIDATES
    stream = gpu.Stream()
    results = []
    for c in cands:
        c = dict(c)
        c.setdefault("stage_precision", precision)
        t0 = time.time()
        try:
            plan = gpu.Plan(mesh, **c)
        except Exception as exc:  # a candidate the mesh cannot use (e.g. LDS too small)
            if log:
                log(f"tune: {c} rejected: {exc}")
            continue
        us = time_plan(plan, dm, precision, iters, stream)
        st = plan.stats()
        del plan
        results.append((c, us, st))
        if log:
            log(f"tune: {c} -> {us:.1f} us (tiles {st['tiles']}, lds {st['lds_bytes']}, "
                f"build+time {time.time() - t0:.1f}s)")
    if not results:
        raise RuntimeError("no tiled-schedule candidate fits this mesh")
    best = min(results, key=lambda r: r[1])
    if save:
        data = {}
        if os.path.exists(TUNED_PATH):
            with open(TUNED_PATH) as f:
                data = json.load(f)
        data[_key(mesh, precision)] = dict(best[0], us=round(best[1], 2))
        with open(TUNED_PATH, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
    return best[0], results

"""Autotuner of the batched step's launch shape (org_tuning,
include/organoid.h) and of the velocity surrogate's GEMM tiles
(include/organoid_fno.h).

Launch shape: the lumen/ECM workgroups of the step's main launch
(lum_blocks), the workgroups of each non-cell role of the launches before
and between steps (roles: psi tiles, h(lum)/h(ECM) parts, lumen/ECM
copy-back), the steps per replayed hipGraph (graph_steps) and where the
lumen/ECM part of the cell update's bracket is formed (shared_bracket: once
per step by the copy-back launch, or by each cell's update).  None of them
changes results (the library's partitions feed only pointwise updates,
per-tile sums in a fixed order and box unions, and both bracket forms
evaluate the same expression on the same operands); they change how the
launches fill the device.  The defaults come from the compute-unit count.

Launch tuning times each candidate either after a warm-up past the slower
first steps that follow an upload (warmup=100, steps=200: the steady
state) or on exactly the steps a caller is about to run (warmup=0, steps=n:
the cache key then carries the horizon), from the same state every time.

GEMM tiles: 64 x 64, 32 x 64, 64 x 32 or 32 x 32 per stage of the network,
chosen by timing forward passes (inference stages) and training steps
(backward stages); every tile accumulates the same products in the same
order, so results do not change either.

Launch tuning times each candidate on the same state (re-uploaded before every
candidate, so each sees the same reshaping work), after a warm-up past the
slower first steps that follow an upload, by
coordinate descent over the three knobs, and keeps the fastest.  Results are
cached in a JSON file keyed by the device and the workload's shape
(grid side, cell count, zone size), so a later run of the same shape on the
same device reuses them; the tile choices likewise per network shape and
batch.
"""
from __future__ import annotations

import json
import os
import time

from . import host

__all__ = ["CACHE_ENV", "default_cache_path", "cache_key", "load_cache", "save_cache", "candidates", "tune",
           "tuned", "tune_fno", "tuned_fno"]

CACHE_ENV = "ORGANOID_TUNE_CACHE"


def default_cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "organoid_mi355x",
                                                     "tuning.json")


def cache_key(device: str, N: int, ncell: int, zone_size: int) -> str:
    return f"{device}|N={N}|cells={ncell}|zone={zone_size}"


def load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(path: str, cache: dict) -> None:
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = f"{path}.{os.getpid()}.tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def candidates(ncu: int) -> dict:
    """Values tried per knob on a device with ncu compute units."""
    return {"lum_blocks": [v for v in (16, 32, 48, 64, 96, 128) if v <= min(128, ncu)],
            "roles": [32, 64, 96, 128, 192, 256],
            "graph_steps": [1, 4, 8, 16, 32],
            "shared_bracket": [1, 0]}


def device_name() -> tuple:
    """(name for the cache key, compute units)."""
    import torch
    p = torch.cuda.get_device_properties(0)
    arch = getattr(p, "gcnArchName", "") or ""
    return f"{p.name}|{arch.split(':')[0]}|cu={p.multi_processor_count}", int(p.multi_processor_count)


def _snapshot(org) -> dict:
    """The state that carries across a step boundary: phi, rho, lum, ECM and
    hfunc (its values outside the current windows are stale and enter psi
    and the areas), so a restored organoid continues as the original."""
    return {f: org.get_field(f) for f in (host.F_PHI, host.F_RHO, host.F_LUM, host.F_ECM, host.F_HFUNC)}


def _restore(org, state: dict) -> None:
    for f, a in state.items():
        org.set_field(f, a)
    org.prepare()


def tune(org, steps: int = 200, warmup: int = 100, passes: int = 2, ncu: int | None = None, log=None) -> dict:
    """Coordinate descent over the knobs from the current tuning, each
    candidate timed on the organoid's current state; leaves the organoid in
    that state with the best tuning set.  Returns {"best", "default",
    "best_ms", "default_ms", "measured": [(tuning, ms), ...]}."""
    if ncu is None:
        ncu = device_name()[1]
    cand = candidates(ncu)
    state = _snapshot(org)
    start = org.get_tuning()
    seen: dict = {}

    def measure(t: dict) -> float:
        key = (t["lum_blocks"], t["roles"], t["graph_steps"], t.get("shared_bracket", 1))
        if key in seen:
            return seen[key]
        org.set_tuning(**t)
        ms = float("inf")
        for _ in range(max(1, 200 // max(steps, 1)) if warmup == 0 else 1):   # short horizons: best of several
            _restore(org, state)
            org.step(warmup, graph=True)
            org.prepare_graph(steps)        # graph capture stays out of the timing
            org.sync()
            t0 = time.perf_counter()
            org.step(steps, graph=True)
            org.sync()
            ms = min(ms, 1e3 * (time.perf_counter() - t0) / steps)
        seen[key] = ms
        if log is not None:
            log(f"tune {t}: {ms:.4f} ms/step")
        return ms

    best = dict(start)
    best_ms = default_ms = measure(best)
    for _ in range(passes):
        changed = False
        for knob in ("lum_blocks", "roles", "graph_steps", "shared_bracket"):
            for v in cand[knob]:
                t = dict(best, **{knob: v})
                ms = measure(t)
                if ms < best_ms:
                    best, best_ms, changed = t, ms, True
        if not changed:
            break
    org.set_tuning(**best)
    _restore(org, state)
    return {"best": best, "default": start, "best_ms": best_ms, "default_ms": default_ms,
            "measured": [({"lum_blocks": k[0], "roles": k[1], "graph_steps": k[2], "shared_bracket": k[3]}, ms)
                         for k, ms in seen.items()]}


def tuned(org, cache_path: str | None = None, retune: bool = False, log=None, **kw) -> dict:
    """Set the organoid's tuning from the cache, or tune and store it.
    Returns {"tuning", "source": "cache" | "tuned", ...tune() results when
    tuned}."""
    path = cache_path or default_cache_path()
    dev, ncu = device_name()
    p = org.params
    key = cache_key(dev, p.N, p.ncell, p.zone_size)
    if kw.get("warmup", 100) == 0:          # tuned on the steps about to run
        key += f"|next={kw.get('steps', 200)}"
    cache = load_cache(path)
    if not retune and key in cache:
        t = cache[key]["tuning"]
        org.set_tuning(**t)
        return {"tuning": t, "source": "cache", "key": key, "path": path}
    r = tune(org, ncu=ncu, log=log, **kw)
    cache = load_cache(path)
    cache[key] = {"tuning": r["best"], "ms_per_step": r["best_ms"], "default_ms_per_step": r["default_ms"]}
    save_cache(path, cache)
    return dict(r, tuning=r["best"], source="tuned", key=key, path=path)


def tune_fno(net, batch: int, reps: int = 5, train: bool = True, passes: int = 2, log=None) -> dict:
    """Coordinate descent over the tile shape of each network stage: the
    forward stages timed by fno_forward, the training stages by
    fno_loss_grad, on random inputs of the given batch.  Sets and returns the
    best tiles ({"tiles", "forward_ms", "train_ms", "default_forward_ms",
    "default_train_ms"})."""
    import torch

    from . import fno as fnomod
    M = net.cfg.M
    x = torch.rand((batch, 2, M, M), device="cuda")
    y = torch.rand((batch, 2, M, M), device="cuda") * 1e-3
    out = torch.empty_like(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn) -> float:
        fn()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    def fwd():
        net.forward(x, out)

    def trn():
        net.loss_grad(x, y)

    tiles = net.get_tiles()
    fwd_stages = [fnomod.STAGES.index(s) for s in ("y", "x", "mix", "ix", "layer")]
    trn_stages = [fnomod.STAGES.index(s) for s in ("proj", "bwd_spectral", "bwd_pointwise", "wgrad")]
    res = {}
    for kind, stages, fn in (("forward", fwd_stages, fwd), ("train", trn_stages if train else [], trn)):
        if not stages:
            continue
        net.set_tiles(tiles)
        best = timed(fn)
        res[f"default_{kind}_ms"] = best
        for _ in range(passes):
            changed = False
            for st in stages:
                for t in range(len(fnomod.TILES)):
                    if t == tiles[st]:
                        continue
                    cand = list(tiles)
                    cand[st] = t
                    net.set_tiles(cand)
                    ms = timed(fn)
                    if log is not None:
                        log(f"tune fno {kind} stage {fnomod.STAGES[st]} tile {fnomod.TILES[t]}: {ms:.4f} ms")
                    if ms < best:
                        best, tiles, changed = ms, cand, True
            if not changed:
                break
        res[f"{kind}_ms"] = best
    net.set_tiles(tiles)
    res["tiles"] = {fnomod.STAGES[k]: fnomod.TILES[v] for k, v in enumerate(tiles)}
    res["tile_ids"] = tiles
    return res


def tuned_fno(net, batch: int, cache_path: str | None = None, retune: bool = False, log=None, **kw) -> dict:
    """Set the network's tiles from the cache, or tune and store them (keyed
    by device, network shape and batch)."""
    path = cache_path or default_cache_path()
    dev, _ = device_name()
    c = net.cfg
    key = f"{dev}|fno M={c.M} w={c.width} m={c.modes} L={c.layers} p={c.proj}|batch={batch}"
    cache = load_cache(path)
    if not retune and key in cache:
        net.set_tiles(cache[key]["tile_ids"])
        return dict(cache[key], source="cache", key=key, path=path)
    r = tune_fno(net, batch, log=log, **kw)
    cache = load_cache(path)
    cache[key] = {k: v for k, v in r.items()}
    save_cache(path, cache)
    return dict(r, source="tuned", key=key, path=path)

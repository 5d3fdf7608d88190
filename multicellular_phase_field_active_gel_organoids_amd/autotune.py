"""Autotuner of the batched step's launch shape (org_tuning, include/organoid.h).

The knobs are the lumen/ECM workgroups of the step's main launch
(lum_blocks), the workgroups of each non-cell role of the launches before
and between steps (roles: psi tiles, h(lum)/h(ECM) parts, lumen/ECM
copy-back) and the steps per replayed hipGraph (graph_steps).  None of them
changes results (the library's partitions feed only pointwise updates,
per-tile sums in a fixed order and box unions); they change how the
launches fill the device.  The defaults come from the compute-unit count.

Tuning times each candidate on the same state (re-uploaded before every
candidate, so each sees the same reshaping work), after a warm-up past the
slower first steps that follow an upload, by
coordinate descent over the three knobs, and keeps the fastest.  Results are
cached in a JSON file keyed by the device and the workload's shape
(grid side, cell count, zone size), so a later run of the same shape on the
same device reuses them.
"""
from __future__ import annotations

import json
import os
import time

from . import host

__all__ = ["CACHE_ENV", "default_cache_path", "cache_key", "load_cache", "save_cache", "candidates", "tune",
           "tuned"]

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
            "graph_steps": [1, 4, 8, 16, 32]}


def device_name() -> tuple:
    """(name for the cache key, compute units)."""
    import torch
    p = torch.cuda.get_device_properties(0)
    arch = getattr(p, "gcnArchName", "") or ""
    return f"{p.name}|{arch.split(':')[0]}|cu={p.multi_processor_count}", int(p.multi_processor_count)


def _snapshot(org) -> dict:
    return {f: org.get_field(f) for f in (host.F_PHI, host.F_RHO, host.F_LUM, host.F_ECM)}


def _restore(org, state: dict) -> None:
    import numpy as np
    for f, a in state.items():
        org.set_field(f, a)
    org.set_field(host.F_HFUNC, np.zeros_like(state[host.F_PHI]))
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
        key = (t["lum_blocks"], t["roles"], t["graph_steps"])
        if key in seen:
            return seen[key]
        org.set_tuning(**t)
        _restore(org, state)
        org.step(warmup, graph=True)
        org.sync()
        t0 = time.perf_counter()
        org.step(steps, graph=True)
        org.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        seen[key] = ms
        if log is not None:
            log(f"tune {t}: {ms:.4f} ms/step")
        return ms

    best = dict(start)
    best_ms = default_ms = measure(best)
    for _ in range(passes):
        changed = False
        for knob in ("lum_blocks", "roles", "graph_steps"):
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
            "measured": [({"lum_blocks": k[0], "roles": k[1], "graph_steps": k[2]}, ms) for k, ms in seen.items()]}


def tuned(org, cache_path: str | None = None, retune: bool = False, log=None, **kw) -> dict:
    """Set the organoid's tuning from the cache, or tune and store it.
    Returns {"tuning", "source": "cache" | "tuned", ...tune() results when
    tuned}."""
    path = cache_path or default_cache_path()
    dev, ncu = device_name()
    p = org.params
    key = cache_key(dev, p.N, p.ncell, p.zone_size)
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

"""Launch tuning of the batched step (org_set_tuning, autotune.py) on the
GPU: every tuning gives the default tuning's trajectory bit for bit (also
when it changes between steps), out-of-range tunings are refused, and the
autotuner leaves the organoid's state as it found it and caches its pick."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from multicellular_phase_field_active_gel_organoids_amd import autotune, host  # noqa: E402
from multicellular_phase_field_active_gel_organoids_amd.configs import make_workload  # noqa: E402

FIELDS = (host.F_PHI, host.F_RHO, host.F_LUM, host.F_ECM)


def _org(wl):
    g = host.Organoid(wl.params, wl.s0_cell)
    g.load_workload(wl)
    return g


def test_tunings_give_the_same_trajectory():
    wl = make_workload("organoid16_512")
    print("case organoid16_512: 24 steps under three tunings", flush=True)
    ref = _org(wl)
    ref.step(24, graph=True)
    for t in ({"lum_blocks": 24, "roles": 40, "graph_steps": 3, "shared_bracket": 1},
              {"lum_blocks": 128, "roles": 256, "graph_steps": 16, "shared_bracket": 1},
              {"lum_blocks": 64, "roles": 96, "graph_steps": 8, "shared_bracket": 0}):
        g = _org(wl)
        g.set_tuning(**t)
        assert g.get_tuning() == t
        g.step(24, graph=True)
        for f in FIELDS:
            assert np.array_equal(g.get_field(f), ref.get_field(f)), (t, f)
    g = _org(wl)                       # tuning changed between steps (new copy-back partition,
    g.step(8, graph=True)              # the bracket's shared plane off, then on again)
    g.set_tuning(lum_blocks=40, roles=72, graph_steps=5, shared_bracket=0)
    g.step(8, graph=True)
    g.set_tuning(lum_blocks=40, roles=72, graph_steps=5, shared_bracket=1)
    g.step(8, graph=True)
    for f in FIELDS:
        assert np.array_equal(g.get_field(f), ref.get_field(f)), f


def test_remainder_graphs_match_eager():
    """A replay shorter than graph_steps, or one that does not divide by it,
    runs its last n % graph_steps steps as one graph of that length (k_step +
    k_mid between steps): 20 steps under 32-step graphs, then 45 = 32 + 13,
    then 20 again (the cached 13-step graph is recaptured), each equal to
    eager steps bit for bit."""
    wl = make_workload("two_cells_256")
    print("case two_cells_256: remainder graphs (20, 45, 20 steps under graph_steps=32)", flush=True)
    g, e = _org(wl), _org(wl)
    t = g.get_tuning()
    g.set_tuning(lum_blocks=t["lum_blocks"], roles=t["roles"], graph_steps=32)
    done = 0
    for n in (20, 45, 20):
        g.step(n, graph=True, it0=done)
        e.step(n, graph=False, it0=done)
        done += n
        for f in FIELDS + (host.F_HFUNC, host.F_PSI):
            assert np.array_equal(g.get_field(f), e.get_field(f)), (n, f)
        assert all((a == b).all() for a, b in zip(g.centres(), e.centres())), n


def test_out_of_range_tunings_are_refused():
    wl = make_workload("two_cells_256")
    g = _org(wl)
    t0 = g.get_tuning()
    for bad in ({"lum_blocks": 0}, {"lum_blocks": 129}, {"roles": 0}, {"roles": 513}, {"graph_steps": 0},
                {"graph_steps": 65}, {"shared_bracket": 2}):
        with pytest.raises(host.OrganoidError):
            g.set_tuning(**dict(t0, **bad))
    assert g.get_tuning() == t0


def test_autotune_restores_state_and_caches(tmp_path):
    wl = make_workload("two_cells_256")
    print("case two_cells_256: autotune, one pass of 8-step timings", flush=True)
    g = _org(wl)
    g.step(4)
    before = {f: g.get_field(f) for f in FIELDS}
    path = str(tmp_path / "tune.json")
    r = autotune.tuned(g, cache_path=path, steps=8, warmup=2, passes=1)
    assert r["source"] == "tuned" and os.path.exists(path)
    assert g.get_tuning() == r["tuning"]
    for f in FIELDS:
        assert np.array_equal(g.get_field(f), before[f]), f
    assert len(r["measured"]) >= 3 and r["best_ms"] <= r["default_ms"]
    g2 = _org(wl)
    r2 = autotune.tuned(g2, cache_path=path)
    assert r2["source"] == "cache" and r2["tuning"] == r["tuning"] == g2.get_tuning()


def test_graph_prepare_leaves_state_and_replays_identically():
    """org_graph_prepare captures the graphs of an n-step replay without
    running them: the state is unchanged (also right after org_prepare,
    whose first step org_step runs eagerly), and the replay that follows
    equals a replay without it bit for bit."""
    wl = make_workload("two_cells_256")
    print("case two_cells_256: org_graph_prepare then 20 + 13 graph steps", flush=True)
    g, r = _org(wl), _org(wl)
    t = g.get_tuning()
    g.set_tuning(lum_blocks=t["lum_blocks"], roles=t["roles"], graph_steps=8)
    r.set_tuning(lum_blocks=t["lum_blocks"], roles=t["roles"], graph_steps=8)
    before = [g.get_field(f) for f in FIELDS + (host.F_HFUNC,)]
    g.prepare_graph(20)
    g.sync()
    for f, a in zip(FIELDS + (host.F_HFUNC,), before):
        assert np.array_equal(g.get_field(f), a), f
    done = 0
    for n in (20, 13):
        g.prepare_graph(n)
        g.step(n, graph=True, it0=done)
        r.step(n, graph=True, it0=done)
        done += n
        for f in FIELDS + (host.F_HFUNC, host.F_PSI):
            assert np.array_equal(g.get_field(f), r.get_field(f)), (n, f)

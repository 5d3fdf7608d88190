"""Batches of organoids (org_batch_*): every member of a batch evolves bit
for bit as it does alone under org_step, whatever the other members' grid
sizes, cell counts and constants (the roles of the shared launches are the
single-organoid roles; only the workgroup counts of the roles that are not
per cell differ, and no result depends on them)."""
import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd import host
from multicellular_phase_field_active_gel_organoids_amd.configs import make_sweep, make_workload

pytestmark = pytest.mark.gpu

FIELDS = (host.F_PHI, host.F_RHO, host.F_LUM, host.F_ECM, host.F_PSI, host.F_HFUNC, host.F_HLUM, host.F_HECM)


def _member(wl):
    g = host.Organoid(wl.params, wl.s0_cell)
    g.load_workload(wl)
    return g


def _same(a, b, what):
    for f in FIELDS:
        assert np.array_equal(a.get_field(f), b.get_field(f)), (what, f)
    assert all((x == y).all() for x, y in zip(a.centres(), b.centres())), what
    ia, ib = a.iterations(), b.iterations()
    assert (ia[0] == ib[0]).all() and ia[1:] == ib[1:], what


def _three():
    w1 = make_workload("two_cells_256")
    w2 = make_workload("hex", N=320, ncell=6)
    w3 = make_workload("hex", N=256, ncell=4)
    w3.params.friction = 2.5
    w3.params.a_ort_ap = w3.params.a_ort_bas = 60.0
    w3.params.b_lat = 0.5
    return [w1, w2, w3]


@pytest.mark.parametrize("graph", [True, False])
def test_batch_members_match_single_runs(graph):
    wls = _three()
    print(f"case batch of 3 (256/2, 320/6, 256/4), 12 steps, graph={graph}", flush=True)
    members = [_member(w) for w in wls]
    with host.Batch(members) as b:
        b.step(12, graph=graph)
        b.sync()
        for w, m in zip(wls, members):
            alone = _member(w)
            alone.step(12, graph=True)
            _same(m, alone, w.name)
        # keep stepping the batch after reading it back, then one member alone
        b.step(5, graph=graph)
    for w, m in zip(wls, members):
        alone = _member(w)
        alone.step(17, graph=True)
        _same(m, alone, w.name + " +5")
    members[0].step(3, graph=True)
    alone = _member(wls[0])
    alone.step(20, graph=True)
    _same(members[0], alone, "member stepped alone after the batch")


def test_sweep8_batch():
    wls = make_sweep()
    print(f"case sweep8: {len(wls)} organoids, {sum(w.params.ncell for w in wls)} cells, 16 steps", flush=True)
    members = [_member(w) for w in wls]
    with host.Batch(members) as b:
        b.step(16, graph=True)
        b.sync()
    for m in members:
        for f in (host.F_PHI, host.F_RHO, host.F_ECM):
            assert np.isfinite(m.get_field(f)).all()
    k = min(range(len(wls)), key=lambda i: wls[i].params.N * wls[i].params.ncell)
    alone = _member(wls[k])
    alone.step(16, graph=True)
    _same(members[k], alone, wls[k].name)

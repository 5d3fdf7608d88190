"""Batches of organoids (org_batch_*): every member of a batch evolves bit
for bit as it does alone under org_step, whatever the other members' grid
sizes, cell counts and constants (the roles of the shared launches are the
single-organoid roles; only the workgroup counts of the roles that are not
per cell differ, and no result depends on them)."""
import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd import host
from multicellular_phase_field_active_gel_organoids_amd.configs import make_sweep, make_workload

from _parity import compare_driver_order, f32_separation, handoff_oracle, rel, round32

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


def _labelled_three():
    """Three organoids of different grid sizes, cell counts and constants,
    each with distinct per-label constants (labels and the viscosity read at
    each mode's pixel matter) and one with a lumen."""
    from multicellular_phase_field_active_gel_organoids_amd.configs import label_constants
    w1 = make_workload("cyst6_256")
    w2 = make_workload("hex", N=320, ncell=7)
    label_constants(w2.params)
    p = w2.params
    p.a_ort_ap, p.a_ort_bas, p.a_ort_lat = 1.7 * p.a_ort_ap, 1.7 * p.a_ort_bas, 1.7 * p.a_ort_lat
    p.b_ap, p.b_bas, p.b_lat = 1.4 * p.b_ap, 1.4 * p.b_bas, 1.4 * p.b_lat
    p.friction = 0.5
    w3 = make_workload("two_cells_256", N=200)
    label_constants(w3.params)
    w3.params.visc_lat, w3.params.deg_bas, w3.params.gen_ap = 3.0, 60.0, 2.0
    return [w1, w2, w3]


def test_batch_members_match_oracle():
    """Each member of a 3-organoid batch (256/6 cyst, 320/7 hex, 200/2, all
    with label-distinct constants) against its own FP64 oracle run over 20
    batched steps (one hipGraph): centres equal, each field's separation at
    most five times the float build's own (plus 1e-6; see tests/_parity.py
    for why label-dependent runs separate), phi / lum / ECM also <= 1e-4.
    Then, from each member's state, one step per function against the
    oracle at the one-step tolerances."""
    wls = [round32(w) for w in _labelled_three()]
    print("case batch of 3 (256/6 cyst, 320/7, 200/2) vs oracle, 20 steps", flush=True)
    members = [_member(w) for w in wls]
    with host.Batch(members) as b:
        b.step(20, graph=True)
        b.sync()
    for w, m in zip(wls, members):
        o, sep = f32_separation(w, 20, fields=("phi", "rho", "ecm", "lum"))
        assert all((x == y).all() for x, y in zip(m.centres(), o.centres())), w.name
        for name, f in (("phi", host.F_PHI), ("rho", host.F_RHO), ("ecm", host.F_ECM), ("lum", host.F_LUM)):
            r = rel(m.get_field(f), o.get(name))
            print(f"  {w.name} {name}: rel L2 {r:.3e} (float build {sep[name]:.3e})", flush=True)
            assert r <= 5.0 * sep[name] + 1e-6, (w.name, name, r, sep[name])
            if name != "rho":
                assert r < 1e-4, (w.name, name)
        o.close()
        compare_driver_order(m, handoff_oracle(m, w), w, f"{w.name} after 20 batched steps")


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

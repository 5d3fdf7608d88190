"""Sweep expansion and the uneven batch workload (host code, no GPU).

create_constant_dictionnary follows main_code/setup_simulation.py:29-50:
every combination of the array-valued constants, the longest array varying
slowest.  Pinned against itertools.product over the keys ordered by
decreasing array length (distinct lengths, so no tie order is involved)."""
import itertools
import math

import numpy as np

from multicellular_phase_field_active_gel_organoids_amd import sweep
from multicellular_phase_field_active_gel_organoids_amd.configs import SWEEP_SIZES, make_sweep
from multicellular_phase_field_active_gel_organoids_amd.params import grid_units


def test_expansion_order_is_the_product_longest_first():
    d = {"a": np.array([1.0, 2.0]), "b": np.array([3.0, 4.0, 5.0]), "c": 7.0, "d": np.array([9.0])}
    got = sweep.create_constant_dictionnary(d)
    want = [{"a": a, "b": bb, "c": 7.0, "d": dd} for bb, a, dd in itertools.product([3.0, 4.0, 5.0], [1.0, 2.0], [9.0])]
    assert len(got) == 6
    assert [{k: float(v) for k, v in g.items()} for g in got] == want


def test_defaults_expand_to_one_run_and_no_arrays_to_none():
    runs = sweep.create_constant_dictionnary(sweep.SWEEP_DEFAULTS)
    assert len(runs) == 1 and runs[0]["ncells"] == 6
    assert sweep.create_constant_dictionnary({"x": 1.0}) == []


def test_constants_to_params_follows_the_template():
    d = dict(sweep.SWEEP_DEFAULTS)
    d.update(ncells=np.array([4, 8]), epsilon_lat=np.array([2.0]), gen_ap=np.array([3.0]), friction=np.array([0.5, 2.0, 4.0]))
    ps = sweep.sweep_params(d, grid_units(256, 4, 20.0))
    assert [(p.ncell, p.friction) for p in ps] == [(4, 0.5), (8, 0.5), (4, 2.0), (8, 2.0), (4, 4.0), (8, 4.0)]
    p = ps[0]
    assert p.a_tan_lat == p.a_ort_lat * 2.0 and p.a_tan_ap == p.a_ort_ap
    assert p.gen_bas == p.gen_ap == 3.0 and p.deg_bas == p.deg_ap == 100.0
    assert p.gen_lat == 1.0 and p.deg_lat == 100.0 and p.n_current == p.ncell


def test_sweep8_workloads():
    a, b = make_sweep(), make_sweep()
    assert len(a) == 8
    base = grid_units(256, 4, 20.0)
    for w, w2 in zip(a, b):
        p = w.params
        assert p.N in SWEEP_SIZES and 4 <= p.ncell <= 64
        assert np.array_equal(w.phi, w2.phi) and np.array_equal(w.rho, w2.rho)
        zs = p.zone_size
        assert np.floor(w.centres).min() - zs - 6 >= 1 and np.ceil(w.centres).max() + zs + 6 <= p.N - 2
        assert 12.0 <= w.radii.min() and w.radii.max() <= 28.0
        for v, v0 in ((p.a_ort_ap, base.a_ort_ap), (p.a_tan_lat, base.a_tan_lat), (p.b_bas, base.b_bas),
                      (p.friction, base.friction)):
            assert abs(math.log10(v / v0)) <= 0.5 + 1e-12
    assert len({w.params.N for w in a}) > 1 and len({w.params.ncell for w in a}) > 1

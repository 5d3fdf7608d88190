"""Helpers shared by the GPU parity tests (not a test module).

Handoff states.  Where the three turnover labels carry different constants,
trajectories of any two implementations separate after a few steps: far
from a cell, phi holds rounding noise (the rho-coupling term of
update_all_cells divides by |grad phi| for any nonzero gradient,
cell_phase_functions.cpp:177-186, and so turns the spectral ringing of rho
into updates of O(1) direction), and the label rule compares phi*lum with
phi*ECM (rho_functions.cpp:328-353), so the *sign* of that noise picks
apical or basal there; through the viscosity read at each mode's pixel
(rho_functions.cpp:149-150) it moves the velocity by ~10 %.  The
reference's own float build and an FP64 build of it separate this way
within 2-3 steps.  So one-step parity is checked from identical float32
states taken along the GPU's own trajectory: from such a state the FP64
oracle's label maps equal the float build's and one call agrees to ~1e-7.
Trajectories are bounded by the float build's own separation from FP64.
"""
from __future__ import annotations

import numpy as np

from multicellular_phase_field_active_gel_organoids_amd import host
from oracle.oracle import Oracle

STATE = (("phi", host.F_PHI), ("rho", host.F_RHO), ("lum", host.F_LUM), ("ecm", host.F_ECM),
         ("hfunc", host.F_HFUNC))


def rel(a, b) -> float:
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def round32(wl):
    """The workload's fields rounded to float32 (the GPU's inputs), so the
    oracle starts from the same values."""
    for k in ("phi", "rho", "lum", "ecm"):
        setattr(wl, k, np.asarray(getattr(wl, k), np.float32).astype(np.float64))
    return wl


def handoff_oracle(g, wl, precision: str = "f64") -> Oracle:
    """An oracle holding the GPU organoid's current state (phi, rho, lum,
    ECM and the stale hfunc outside the windows), with the driver's
    pre-loop copies (run_with_active_gels.cpp:92-98)."""
    o = Oracle(wl.params, wl.s0_cell, precision=precision)
    for name, f in STATE:
        o.set(name, g.get_field(f))
    o.call("init_like_driver")
    return o


def windows(plane, xc, yc, zs, off, side):
    return np.stack([plane[n, xc[n] - zs + off: xc[n] - zs + off + side, yc[n] - zs + off: yc[n] - zs + off + side]
                     for n in range(plane.shape[0])])


def iterations_agree(gi, oi, what):
    """Reshaping counts are threshold decisions (stop when the deviation
    falls below a fixed threshold, cell_phase_functions.cpp:473-475, 774):
    from identical states the FP32 and FP64 sums land on opposite sides of
    it in ~6 % of the steps of cyst6_256 (one cell, one sweep).  Returns the
    number of cells whose count differs; asserts the lumen/ECM counts are
    equal and no cell differs by more than one sweep."""
    assert gi[1:] == oi[1:], (what, gi, oi)
    d = np.abs(np.asarray(gi[0]) - np.asarray(oi[0]))
    assert d.max() <= 1, (what, gi, oi)
    return int((d > 0).sum())


def compare_driver_order(g, o, wl, tag: str, scale: float = 1.0) -> None:
    """One timestep through the per-function entry points on the GPU and the
    oracle, in the driver's order (run_with_active_gels.cpp:123-182), each
    output compared after its call.  Tolerances (relative L2, FP32 vs FP64):
    stencil / pointwise fields 2e-6, spectral outputs 2e-5 (four FP32
    transforms), areas 1e-6; centres and lumen/ECM counts exact.  `scale`
    multiplies every tolerance (1e-5 for an FP64 context, which must also
    match every reshaping count)."""
    zs, W, M = wl.params.zone_size, wl.params.W, wl.params.M
    t = lambda v: v * scale  # noqa: E731
    print(f"case {tag}: per-function step", flush=True)
    g.find_centres_cells(); o.call("find_centres_cells")
    xc, yc = g.centres()
    oxc, oyc = o.centres()
    assert (xc == oxc).all() and (yc == oyc).all(), tag

    o.call("create_turnover_rates_tensions")
    g.compute_hfunc_3d_tables_kernels_loop_n(); o.call("compute_hfunc_cells")
    assert rel(g.get_field(host.F_HFUNC), o.get("hfunc")) < t(2e-6), tag
    g.compute_all_hfunc_2d_tables_kernels(); o.call("compute_hfunc_lum_ecm")
    assert rel(g.get_field(host.F_HECM), o.get("hecm")) < t(2e-6), tag
    assert rel(g.get_field(host.F_HLUM), o.get("hlum")) < t(2e-6), tag
    o.call("compute_normal_grad_phi"); o.call("compute_prod_norm_delta")
    g.compute_psi_kernels(); o.call("compute_psi")
    assert rel(g.get_field(host.F_PSI), o.get("psi")) < t(2e-6), tag
    s = g.sum_s_kernels(); o.call("sum_s")
    os_, of0 = o.areas()
    assert np.allclose(s, os_, rtol=t(1e-6)), tag
    _, f0 = g.areas()
    assert np.allclose(f0, of0, rtol=t(1e-4), atol=t(1e-3)), tag

    g.update_all_cortices(); o.call("update_all_cortices")
    r = {k: rel(g.get_window(w), o.get(k)) for k, w in (("vx", host.W_VX), ("vy", host.W_VY))}
    r["rhon"] = rel(g.get_window(host.W_RHON), windows(o.get("rhon"), xc, yc, zs, 1, M))
    print(f"  spectral: " + ", ".join(f"{k} {v:.2e}" for k, v in r.items()), flush=True)
    assert max(r.values()) < t(2e-5), (tag, r)

    g.update_lumen_and_ECM(); o.call("update_lumen_and_ECM")
    g.reshaping_lumen_or_ECM(0); o.call("reshaping_lumen_or_ECM", 0)
    g.reshaping_lumen_or_ECM(1); o.call("reshaping_lumen_or_ECM", 1)
    assert rel(g.get_field(host.F_ECMN), o.get("ecmn")) < t(2e-6), tag
    assert rel(g.get_field(host.F_LUMN), o.get("lumn")) < t(2e-6), tag

    o.call("compute_normal_to_phi")
    g.update_all_cells(); o.call("update_all_cells")
    assert rel(g.get_window(host.W_PHIN), windows(o.get("phin"), xc, yc, zs, 0, W)) < t(2e-6), tag
    g.reshaping_cells(); o.call("reshaping_cells")
    if scale < 1.0:
        gi, oi = g.iterations(), o.iterations()
        assert gi[1:] == oi[1:] and (gi[0] == oi[0]).all(), (tag, gi, oi)
        ndiff = 0
    else:
        ndiff = iterations_agree(g.iterations(), o.iterations(), tag)
    # one extra sweep on one cell moves its phi by ~1e-5 of the whole field
    assert rel(g.get_field(host.F_PHI), o.get("phi")) < (t(2e-6) if ndiff == 0 else 5e-5), tag
    g.copy_tables(); o.call("commit_step")
    assert rel(g.get_field(host.F_RHO), o.get("rho")) < t(2e-5), tag
    assert rel(g.get_field(host.F_ECM), o.get("ecm")) < t(2e-6), tag
    assert rel(g.get_field(host.F_LUM), o.get("lum")) < t(2e-6), tag


def f32_separation(wl, steps: int, fields=("phi", "rho", "ecm", "lum", "vx", "vy")):
    """The FP64 oracle and the oracle's float build (the reference's own
    arithmetic) run `steps` steps from the same state; returns the FP64
    oracle and the float build's relative L2 separation per field."""
    a = Oracle(wl.params, wl.s0_cell)
    a.load_workload(wl)
    b = Oracle(wl.params, wl.s0_cell, precision="f32")
    b.load_workload(wl)
    for it in range(steps):
        a.step(it)
        b.step(it)
    sep = {k: rel(b.get(k), a.get(k)) for k in fields}
    b.close()
    return a, sep

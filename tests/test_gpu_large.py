"""The largest single-GPU workload, organoid256_2048 (256 cells, radii
U(14, 26) on a jittered hex packing, 2048^2): too large for the FP64 oracle
(its state alone is ~100 GiB), so it is checked through properties that do
not depend on size: the graph-replayed and the eager batched step agree bit
for bit, every field stays finite and in range, every cell keeps its area
within the reshaping tolerance of its target and stays near its centre:
over 24 steps, and over the configuration's full 2 x 10^4 steps (~20 s)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from multicellular_phase_field_active_gel_organoids_amd import host  # noqa: E402
from multicellular_phase_field_active_gel_organoids_amd.configs import make_workload  # noqa: E402


def test_organoid256_2048_steps():
    wl = make_workload("organoid256_2048")
    p = wl.params
    print(f"case organoid256_2048: {p.ncell} cells, {p.N}^2, 24 steps graph vs eager", flush=True)
    g1 = host.Organoid(p, wl.s0_cell)
    g1.load_workload(wl)
    g1.step(24, graph=True)
    g2 = host.Organoid(p, wl.s0_cell)
    g2.load_workload(wl)
    g2.step(24, graph=False)
    for f in (host.F_PHI, host.F_RHO, host.F_LUM, host.F_ECM):
        a = g1.get_field(f)
        assert np.array_equal(a, g2.get_field(f)), f
        assert np.isfinite(a).all(), f
        if f != host.F_RHO:
            assert a.min() > -0.05 and a.max() < 1.05, (f, a.min(), a.max())
    g2.close()
    s, _ = g1.areas()
    rel = np.abs(s - wl.s0_cell) / wl.s0_cell
    print(f"  area / target: max deviation {rel.max():.3f}", flush=True)
    assert rel.max() < 0.1
    xc, yc = g1.centres()
    d = np.hypot(xc - wl.centres[:, 0], yc - wl.centres[:, 1])
    print(f"  centre displacement: max {d.max():.2f} px", flush=True)
    assert d.max() < 4.0


def test_organoid256_2048_full_length():
    """The configuration's full length, 2 x 10^4 steps (replayed as graphs):
    at every 5000 steps every field finite and in range, every cell's area
    near its target and its centre near where it started, the reshaping
    loops within their iteration limit (a limit reached raises)."""
    wl = make_workload("organoid256_2048")
    p = wl.params
    print(f"case organoid256_2048: {p.ncell} cells, {p.N}^2, 2 x 10^4 steps in graphs", flush=True)
    g = host.Organoid(p, wl.s0_cell)
    g.load_workload(wl)
    done = 0
    for chunk in (5000, 5000, 5000, 5000):
        g.step(chunk, graph=True, it0=done)
        done += chunk
        fields = {f: g.get_field(f) for f in (host.F_PHI, host.F_RHO, host.F_LUM, host.F_ECM)}
        ranges = {f: (float(a.min()), float(a.max())) for f, a in fields.items()}
        s, _ = g.areas()
        rel = np.abs(s - wl.s0_cell) / wl.s0_cell
        xc, yc = g.centres()
        d = np.hypot(xc - wl.centres[:, 0], yc - wl.centres[:, 1])
        cells, lum, ecm = g.iterations()
        print(f"  step {done}: ranges phi {ranges[host.F_PHI]} rho {ranges[host.F_RHO]} lum {ranges[host.F_LUM]} "
              f"ecm {ranges[host.F_ECM]}; area dev max {rel.max():.3f}; centre displacement max {d.max():.2f} px; "
              f"sweeps cells <= {int(cells.max())}, lum {lum}, ECM {ecm}", flush=True)
        for f, a in fields.items():
            assert np.isfinite(a).all(), (done, f)
            if f != host.F_RHO:
                assert ranges[f][0] > -0.05 and ranges[f][1] < 1.05, (done, f, ranges[f])
        assert rel.max() < 0.02, (done, rel.max())         # observed 0.002 at every checkpoint
        assert d.max() < 4.0, (done, d.max())              # observed 1.27 px
    g.close()

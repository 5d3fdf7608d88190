"""The reference template's own configuration on the GPU: template6_1200
(the constants of constants_template.cpp with setup_simulation.py's
defaults, 6 cells, dx = 0.01, zone_size 200: FFT box 398 = 2 x 199,
Bluestein transforms; 1200^2) from its IC program's analytic state.  The
first step does the IC program's relaxation (~6700 reshaping iterations per
cell, ~4400 / ~7200 lumen / ECM sweeps), the next ones 80-200 iterations.

Against tests/golden/template6_1200.npz (the FP64 oracle and its float
build, tests/golden/make_template_fixture.py): per step the centres, every
reshaping count and the areas; after the last step the fields on a stride-20
grid and per-cell moments of phi and rho over the whole plane.

* FP64 context: counts and centres equal to the FP64 oracle's, areas within
  1e-12, fields and moments within 1e-9 relative.
* FP32 context (the reference's float build on the full-plane path): counts
  and centres equal to the float build's, fields and moments within 1e-6 of
  it.
"""
import os

import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd import host
from multicellular_phase_field_active_gel_organoids_amd.configs import make_workload

from _parity import rel, round32

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "template6_1200.npz")


def moments(phi, rho):
    N = phi.shape[-1]
    i = np.arange(N, dtype=np.float64)[:, None]
    j = np.arange(N, dtype=np.float64)[None, :]
    return np.stack([phi.sum(axis=(1, 2)), (phi * phi).sum(axis=(1, 2)), (phi * i).sum(axis=(1, 2)),
                     (phi * j).sum(axis=(1, 2)), rho.sum(axis=(1, 2)), (rho * rho).sum(axis=(1, 2))], axis=1)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_template6_1200(precision):
    fx = np.load(GOLDEN)
    steps, stride = int(fx["steps"]), int(fx["stride"])
    wl = make_workload("template6_1200")
    if precision == "f32":
        round32(wl)
    p = precision + "_"
    tol = 1e-9 if precision == "f64" else 1e-6
    g = host.Organoid(wl.params, wl.s0_cell, precision=precision)
    assert g.generic
    g.load_workload(wl)
    for it in range(steps):
        print(f"case template6_1200 {precision}: step {it}", flush=True)
        g.step(1)
        xc, yc = g.centres()
        cells, lum, ecm = g.iterations()
        print(f"  counts {cells} {lum} {ecm} (oracle {fx[p + 'it_cells'][it]} {fx[p + 'it_lum'][it]} "
              f"{fx[p + 'it_ecm'][it]})", flush=True)
        assert (xc == fx[p + "centres"][it, :, 0]).all() and (yc == fx[p + "centres"][it, :, 1]).all()
        assert (cells == fx[p + "it_cells"][it]).all()
        assert lum == fx[p + "it_lum"][it] and ecm == fx[p + "it_ecm"][it]
        s, _ = g.areas()
        assert np.allclose(s, fx[p + "areas"][it], rtol=1e-12 if precision == "f64" else 1e-6)
    phi = g.get_field(host.F_PHI).astype(np.float64)
    rho = g.get_field(host.F_RHO).astype(np.float64)
    r = rel(moments(phi, rho), fx[p + "moments"])
    print(f"  moments rel {r:.2e}", flush=True)
    assert r <= tol
    for name, f in (("phi", host.F_PHI), ("rho", host.F_RHO), ("lum", host.F_LUM), ("ecm", host.F_ECM)):
        got = g.get_field(f).astype(np.float64)[..., ::stride, ::stride]
        r = rel(got, fx[p + name])
        print(f"  {name}: rel L2 on the stride-{stride} grid {r:.2e}", flush=True)
        assert r <= tol, (name, r)
    g.close()

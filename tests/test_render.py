"""PNG renders of a state (render.py, plot_final_steps.py:100-221): the three
images' values on hand-built states, the colour maps, the PNG encoding
(decoded independently by Pillow) and a run directory in the reference's
.dat layout rendered end to end."""
import os

import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd import render, snapshots


def test_phases_labels_priority_and_threshold():
    n = 8
    lum = np.zeros((n, n)); ecm = np.zeros((n, n)); phi = np.zeros((2, n, n))
    lum[0, 0:3] = 1.0            # lumen only at (0,0); lumen+ECM at (0,1); lumen+ECM+cell at (0,2)
    ecm[0, 1:4] = 1.0            # ECM only at (0,3)
    phi[0, 0, 2] = 1.0
    phi[1, 1, 0] = 0.5           # h(0.5) = 0.5: counts (>=)
    phi[1, 1, 1] = 0.4999        # just below
    ecm[1, 2] = 0.5
    lum[1, 3] = np.nextafter(0.5, 0.0)
    img = render.phases_image(phi, lum, ecm)
    want = np.full((n, n), 0.1)
    want[0, 0] = 0.4             # lumen
    want[0, 1] = 1.0             # ECM drawn over lumen
    want[0, 2] = 0.6             # cell drawn over both
    want[0, 3] = 1.0
    want[1, 0] = 0.6
    want[1, 2] = 1.0
    np.testing.assert_array_equal(img, want)


def test_cells_and_cortices_images():
    phi = np.zeros((2, 2, 3))
    phi[0] = [[0.3, 0.5, 0.8], [1.0, 0.49, 0.0]]
    phi[1] = [[0.6, 0.6, 0.1], [0.0, 0.7, 0.9]]
    np.testing.assert_allclose(render.cells_image(phi), [[0.6, 1.1, 0.8], [1.0, 0.7, 0.9]])
    rho = np.zeros((2, 2, 2))
    rho[0] = [[-1.0, 2.0], [0.5, -3.0]]
    rho[1] = [[-0.5, 1.0], [0.7, -0.1]]
    np.testing.assert_array_equal(render.cortices_image(rho), [[0.0, 2.0], [0.7, 0.0]])


def test_colour_maps():
    # the four bands of the phase map: 64 of 256 entries each, as imshow indexes them
    v = np.array([0.0, 0.1, 0.249, 0.25, 0.4, 0.499, 0.5, 0.6, 0.75, 1.0])
    rgb = render.colourize(v, "phases")
    bands = [0, 0, 0, 1, 1, 1, 2, 2, 3, 3]
    np.testing.assert_array_equal(rgb, np.round(render.PHASE_COLOURS[bands] * 255).astype(np.uint8))
    assert tuple(rgb[1]) == (0, 0, 0) and tuple(rgb[-1]) == (255, 235, 13)
    # continuous maps: scaled to min..max, ends at the anchors' ends, constant image -> first colour
    x = np.linspace(-2.0, 5.0, 101)
    for name in ("CMRmap", "viridis"):
        c = render.colourize(x, name).astype(int)
        a = np.round(render.MAPS[name] * 255).astype(int)
        np.testing.assert_array_equal(c[0], a[0]); np.testing.assert_array_equal(c[-1], a[-1])
        np.testing.assert_array_equal(render.colourize(np.full(4, 3.0), name)[0], a[0])
    lum = render.colourize(x, "viridis").astype(float) @ [0.2126, 0.7152, 0.0722]
    assert (np.diff(lum) >= -1.0).all()      # viridis brightens monotonically


def test_png_round_trip_and_independent_decoder(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(5)
    rgb = rng.integers(0, 256, (13, 7, 3), dtype=np.uint8)
    p = str(tmp_path / "a.png")
    render.write_png(p, rgb)
    np.testing.assert_array_equal(render.read_png(p), rgb)
    with Image.open(p) as im:
        assert im.mode == "RGB" and im.size == (7, 13)
        np.testing.assert_array_equal(np.asarray(im), rgb)
    render.write_png(p, rgb, scale=3)
    big = render.read_png(p)
    assert big.shape == (39, 21, 3)
    np.testing.assert_array_equal(big[::3, ::3], rgb)
    b = bytearray(open(p, "rb").read())
    b[40] ^= 0xFF                            # inside IDAT
    open(p, "wb").write(bytes(b))
    with pytest.raises(ValueError, match="CRC"):
        render.read_png(p)


def test_render_run_directory(tmp_path):
    """A run directory written in the reference's layout (two frames, two
    cells), rendered from its last .dat frame into all_plots/final_step/."""
    Image = pytest.importorskip("PIL.Image")
    N, nc, nf = 24, 2, 2
    yy, xx = np.mgrid[0:N, 0:N]
    phi = np.zeros((nf, nc, N, N), np.float32)
    rho = np.zeros((nf, nc, N, N), np.float32)
    for f in range(nf):
        for n in range(nc):
            r = np.hypot(yy - 12, xx - (7 + 10 * n) - f)
            phi[f, n] = 0.5 * (1 - np.tanh((r - 4.0) / 1.5))
            rho[f, n] = np.where((phi[f, n] > 0.1) & (phi[f, n] < 0.9), 1.0 + 0.1 * n + f, 0.0)
    lum = np.zeros((nf, N, N), np.float32)
    ecm = np.zeros((nf, N, N), np.float32)
    ecm[:, :, :] = (np.hypot(yy - 12, xx - 12) > 10).astype(np.float32)
    lum[:, 12, 12] = 1.0
    slabs = {"phiout": phi, "rhoout": rho, "lumout": lum, "ECMout": ecm}
    snapshots.write_run_outputs(str(tmp_path), slabs, nsaving=50)
    out = render.render_run(str(tmp_path), N, nc)
    assert out["step"] == 50
    d = tmp_path / "all_plots" / "final_step"
    assert sorted(os.listdir(d)) == ["all_cells_final.png", "all_cortices_final.png", "all_phases.png"]
    last = {k: snapshots.read_dat(str(tmp_path / "data_int" / f), N)
            for k, f in (("lum", "lumen_step_50.dat"), ("ecm", "ECM_step_50.dat"))}
    ph = np.stack([snapshots.read_dat(str(tmp_path / "data_int" / f"phi_cell_{n + 1}_step_50.dat"), N)
                   for n in range(nc)])
    with Image.open(d / "all_phases.png") as im:
        got = np.asarray(im)
    np.testing.assert_array_equal(got, render.colourize(render.phases_image(ph, last["lum"], last["ecm"])))
    # the corners are ECM (yellow), a cell's centre is red
    assert tuple(got[0, 0]) == (255, 235, 13)
    assert tuple(got[12, 7 + 1]) == (250, 5, 0)

"""PNG renders of an organoid state, for visual regression: the three figures
of the reference's final-step plot (plot_final_steps.py:100-221), without
matplotlib, one image pixel per grid point (optionally scaled up).

* cortices -- pixelwise maximum of the cells' rho over a zero background
  (plot_final_steps.py:100-114), colour map CMRmap scaled to the image's
  min..max (:118).
* cells -- sum over cells of phi where phi >= 0.5, else 0
  (plot_final_steps.py:193-199), colour map viridis scaled to min..max (:204).
* phases -- 0.4 where h(lum) >= 0.5, then 1.0 where h(ECM) >= 0.5, then 0.6
  where h(phi_n) >= 0.5 for any cell, then 0.1 where still below 0.4, with
  h(u) = u^2 (3 - 2u) (plot_final_steps.py:172-203); drawn with the
  reference's four-band colour map (black, blue, red, yellow; :149-166).

``render_run`` reads the last saved frame of a run directory in the
reference's ``.dat`` layout (``data_int/phi_cell_<n>_step_<s>.dat`` etc.,
as ``snapshots.write_run_outputs`` writes it) and writes
``all_plots/final_step/{all_cortices_final,all_cells_final,all_phases}.png``
as the reference's script does.
"""
from __future__ import annotations

import os
import re
import struct
import zlib

import numpy as np

from . import snapshots

__all__ = ["cortices_image", "cells_image", "phases_image", "colourize", "write_png", "read_png",
           "render_state", "render_run", "PHASE_COLOURS"]

# plot_final_steps.py:149-166: the four quarters of a 256-entry map over 0..1
PHASE_COLOURS = np.array([[0.0, 0.0, 0.0],
                          [0.0784313725490196, 0.6, 0.9803921568627451],
                          [0.9803921568627451, 0.0196078431372549, 0.0],
                          [1.0, 0.9215686274509803, 0.050980392156862744]])

# anchor colours of the two continuous maps (matplotlib's CMRmap: nine
# equally spaced anchors; viridis: five samples at 0, 1/4, 1/2, 3/4, 1),
# interpolated linearly between anchors
_CMRMAP = np.array([[0.0, 0.0, 0.0], [0.15, 0.15, 0.5], [0.3, 0.15, 0.75], [0.6, 0.2, 0.5],
                    [1.0, 0.25, 0.15], [0.9, 0.5, 0.0], [0.9, 0.75, 0.1], [0.9, 0.9, 0.5], [1.0, 1.0, 1.0]])
_VIRIDIS = np.array([[0x44, 0x01, 0x54], [0x3b, 0x52, 0x8b], [0x21, 0x91, 0x8c], [0x5e, 0xc9, 0x62],
                     [0xfd, 0xe7, 0x25]]) / 255.0
MAPS = {"CMRmap": _CMRMAP, "viridis": _VIRIDIS}


def _h(u):
    u = np.asarray(u, np.float64)
    return u * u * (3.0 - 2.0 * u)


def cortices_image(rho: np.ndarray) -> np.ndarray:
    """Pixelwise max over cells of rho, starting from 0 (plot_final_steps.py:
    103-114: a value replaces the image only where it is larger)."""
    rho = np.asarray(rho, np.float64)
    return np.maximum(rho.max(axis=0), 0.0)


def cells_image(phi: np.ndarray) -> np.ndarray:
    """Sum over cells of phi with the values below 0.5 set to 0
    (plot_final_steps.py:197-199)."""
    phi = np.asarray(phi, np.float64)
    return np.where(phi < 0.5, 0.0, phi).sum(axis=0)


def phases_image(phi: np.ndarray, lum: np.ndarray, ecm: np.ndarray) -> np.ndarray:
    """Phase labels 0.1 / 0.4 / 0.6 / 1.0 (plot_final_steps.py:177-203):
    lumen, then ECM, then cells overwrite a zero image; what stays below 0.4
    becomes 0.1."""
    img = np.zeros(np.asarray(lum).shape)
    img[_h(lum) >= 0.5] = 0.4
    img[_h(ecm) >= 0.5] = 1.0
    for p in np.asarray(phi):
        img[_h(p) >= 0.5] = 0.6
    img[img < 0.4] = 0.1
    return img


def colourize(img: np.ndarray, cmap: str = "phases", vmin: float | None = None,
              vmax: float | None = None) -> np.ndarray:
    """RGB uint8 image.  ``phases``: the four-band map over 0..1 (256 entries,
    64 per colour, nearest entry as imshow with vmin 0, vmax 1); ``CMRmap`` /
    ``viridis``: scaled to vmin..vmax (default the image's min..max)."""
    img = np.asarray(img, np.float64)
    if cmap == "phases":
        idx = np.clip(np.floor(np.clip(img, 0.0, 1.0) * 256.0), 0, 255).astype(np.int64) // 64
        rgb = PHASE_COLOURS[idx]
    else:
        anchors = MAPS[cmap]
        lo = float(np.nanmin(img)) if vmin is None else vmin
        hi = float(np.nanmax(img)) if vmax is None else vmax
        t = np.zeros_like(img) if hi <= lo else np.clip((img - lo) / (hi - lo), 0.0, 1.0)
        x = t * (len(anchors) - 1)
        k = np.minimum(np.floor(x).astype(np.int64), len(anchors) - 2)
        f = (x - k)[..., None]
        rgb = anchors[k] * (1.0 - f) + anchors[k + 1] * f
    return np.round(rgb * 255.0).astype(np.uint8)


def _chunk(kind: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + kind + data + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def write_png(path: str, rgb: np.ndarray, scale: int = 1) -> None:
    """8-bit RGB PNG (no interlace, filter type 0 on every row), each pixel
    repeated scale x scale times."""
    rgb = np.asarray(rgb, np.uint8)
    if scale > 1:
        rgb = np.repeat(np.repeat(rgb, scale, axis=0), scale, axis=1)
    h, w = rgb.shape[:2]
    raw = np.concatenate([np.zeros((h, 1), np.uint8), rgb.reshape(h, w * 3)], axis=1).tobytes()
    png = (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0))
           + _chunk(b"IDAT", zlib.compress(raw, 6)) + _chunk(b"IEND", b""))
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(png)


def read_png(path: str) -> np.ndarray:
    """Read back a PNG of the form write_png produces (8-bit RGB, filter 0),
    checking the signature and every chunk's CRC."""
    with open(path, "rb") as f:
        b = f.read()
    if b[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos, idat, w, h = 8, b"", None, None
    while pos < len(b):
        (n,) = struct.unpack(">I", b[pos:pos + 4])
        kind, data = b[pos + 4:pos + 8], b[pos + 8:pos + 8 + n]
        (crc,) = struct.unpack(">I", b[pos + 8 + n:pos + 12 + n])
        if crc != zlib.crc32(kind + data) & 0xFFFFFFFF:
            raise ValueError(f"bad CRC in chunk {kind!r}")
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, inter = struct.unpack(">IIBBBBB", data)
            if (depth, ctype, inter) != (8, 2, 0):
                raise ValueError("only 8-bit RGB, non-interlaced PNGs are read")
        elif kind == b"IDAT":
            idat += data
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    if (raw[:, 0] != 0).any():
        raise ValueError("only filter type 0 is read")
    return raw[:, 1:].reshape(h, w, 3).copy()


def render_state(outdir: str, phi: np.ndarray, rho: np.ndarray, lum: np.ndarray, ecm: np.ndarray,
                 scale: int = 1) -> dict:
    """The three final-step figures of a state into outdir; returns their
    paths."""
    out = {"cortices": os.path.join(outdir, "all_cortices_final.png"),
           "cells": os.path.join(outdir, "all_cells_final.png"),
           "phases": os.path.join(outdir, "all_phases.png")}
    write_png(out["cortices"], colourize(cortices_image(rho), "CMRmap"), scale)
    write_png(out["cells"], colourize(cells_image(phi), "viridis"), scale)
    write_png(out["phases"], colourize(phases_image(phi, lum, ecm), "phases"), scale)
    return out


def render_run(rundir: str, N: int, ncell: int, scale: int = 1) -> dict:
    """Render the last saved frame of a run directory (``data_int/*.dat``)
    into ``rundir/all_plots/final_step/`` (plot_final_steps.py:116-221)."""
    dd = os.path.join(rundir, "data_int")
    steps = sorted({int(m.group(1)) for f in os.listdir(dd) for m in [re.match(r"lumen_step_(\d+)\.dat$", f)] if m})
    if not steps:
        raise FileNotFoundError(f"no lumen_step_<s>.dat in {dd}")
    s = steps[-1]
    phi = np.stack([snapshots.read_dat(os.path.join(dd, f"phi_cell_{n + 1}_step_{s}.dat"), N) for n in range(ncell)])
    rho = np.stack([snapshots.read_dat(os.path.join(dd, f"rho_cell_{n + 1}_step_{s}.dat"), N) for n in range(ncell)])
    lum = snapshots.read_dat(os.path.join(dd, f"lumen_step_{s}.dat"), N)
    ecm = snapshots.read_dat(os.path.join(dd, f"ECM_step_{s}.dat"), N)
    out = render_state(os.path.join(rundir, "all_plots", "final_step"), phi, rho, lum, ecm, scale)
    out["step"] = s
    return out

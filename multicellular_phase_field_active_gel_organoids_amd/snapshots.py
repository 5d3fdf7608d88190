"""State files and snapshots in the reference's on-disk formats, and the
driver's saving schedule.

Formats (all planes row-major, i = row):

* ``.bin`` -- the state files a run starts from and leaves behind
  (``ini_cell_<n>.bin``, ``ini_rho_<n>.bin``, ``ini_lumen.bin``,
  ``ini_ECM.bin``).  The reference writes every element as
  ``sizeof(float*)`` = 8 bytes read from the element's address
  (run_with_active_gels.cpp:197-201) and reads it back the same way
  (initialisation_functions.cpp:16-18): element k sits at byte 8k and bytes
  8k+4..8k+7 hold the next float of the writer's slab (for the slab's last
  element the reference reads past its allocation; 0 is written here).
* ``.dat`` -- text snapshots: one value per line printed with ``%.15f`` from
  the float promoted to double, and a blank line after each row
  (run_with_active_gels.cpp:215-267).

Saving schedule: the driver saves the state at the *start* of every step
``it`` with ``it % nsaving == 0`` (run_with_active_gels.cpp:110-116), so a run
of ``Nsteps`` steps keeps ``1 + (Nsteps - 1) // nsaving`` frames and the
result of the last step is saved only if it is itself a frame.  After the
loop every frame is written in order and the ``.bin`` files are overwritten
each time (run_with_active_gels.cpp:191-269): they end up holding the last
saved frame, which is what a restart reads.
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["write_bin", "read_bin", "write_dat", "read_dat", "frame_count", "run_saving",
           "write_run_outputs", "read_state_files"]


def write_bin(path: str, slab: np.ndarray, start: int = 0, count: int | None = None) -> None:
    """Write ``slab[start:start+count]`` (float32, flattened) as 8-byte records
    (value, next value of the slab)."""
    s = np.ascontiguousarray(slab, np.float32).ravel()
    if count is None:
        count = s.size - start
    if start < 0 or count < 0 or start + count > s.size:
        raise ValueError(f"range [{start}, {start + count}) outside a slab of {s.size} floats")
    rec = np.zeros((count, 2), np.float32)
    rec[:, 0] = s[start:start + count]
    tail = min(count, s.size - start - 1)
    if tail > 0:
        rec[:tail, 1] = s[start + 1:start + 1 + tail]
    with open(path, "wb") as f:
        f.write(rec.tobytes())


def read_bin(path: str, count: int) -> np.ndarray:
    """``count`` floats of an 8-byte-record ``.bin`` file (first 4 bytes of
    each record); raises if the file is shorter."""
    with open(path, "rb") as f:
        data = f.read(8 * count)
    if len(data) != 8 * count:
        raise ValueError(f"{path}: {len(data)} bytes, expected {8 * count}")
    return np.frombuffer(data, np.float32)[0::2].copy()


def write_dat(path: str, plane: np.ndarray, fmt: str = "%.15f") -> None:
    """One value per line, a blank line after each row."""
    a = np.asarray(plane, np.float32)
    if a.ndim != 2:
        raise ValueError("write_dat takes one plane")
    with open(path, "w") as f:
        for row in a:
            f.write("".join((fmt % float(v)) + "\n" for v in row))
            f.write("\n")


def read_dat(path: str, N: int) -> np.ndarray:
    """An N x N plane from a ``.dat`` snapshot (blank lines ignored)."""
    with open(path) as f:
        vals = [float(t) for t in f.read().split()]
    if len(vals) != N * N:
        raise ValueError(f"{path}: {len(vals)} values, expected {N * N}")
    return np.array(vals, np.float64).astype(np.float32).reshape(N, N)


def frame_count(Nsteps: int, nsaving: int) -> int:
    """Frames a run of Nsteps keeps (constants_template.cpp:14)."""
    return 1 + (Nsteps - 1) // nsaving


def run_saving(org, Nsteps: int, nsaving: int, graph: bool = True) -> dict:
    """Advance ``org`` (a host.Organoid) by ``Nsteps`` steps, saving phi, rho,
    lum and ECM at the start of every step ``it % nsaving == 0``
    (run_with_active_gels.cpp:102-116).  The steps between two frames run as
    one batched call.  Returns the slabs ``phiout`` (frames, ncell, N, N),
    ``rhoout``, ``lumout`` and ``ECMout`` (frames, N, N)."""
    from . import host
    if Nsteps < 1 or nsaving < 1:
        raise ValueError("Nsteps and nsaving must be positive")
    nf = frame_count(Nsteps, nsaving)
    out = {"phiout": np.empty((nf, org.nc, org.N, org.N), np.float32),
           "rhoout": np.empty((nf, org.nc, org.N, org.N), np.float32),
           "lumout": np.empty((nf, org.N, org.N), np.float32),
           "ECMout": np.empty((nf, org.N, org.N), np.float32)}
    it = 0
    for fr in range(nf):
        out["phiout"][fr] = org.get_field(host.F_PHI)
        out["rhoout"][fr] = org.get_field(host.F_RHO)
        out["lumout"][fr] = org.get_field(host.F_LUM)
        out["ECMout"][fr] = org.get_field(host.F_ECM)
        n = min(nsaving, Nsteps - it)
        org.step(n, graph=graph, it0=it)
        it += n
    return out


def write_run_outputs(outdir: str, slabs: dict, nsaving: int, n_current: int | None = None) -> None:
    """The driver's post-loop writes (run_with_active_gels.cpp:191-269), in
    its order: per frame, per cell ``ini_cell_<n>.bin``, ``ini_rho_<n>.bin``,
    ``data_int/phi_cell_<n>_step_<s>.dat``, ``data_int/rho_cell_<n>_step_<s>.dat``;
    then ``ini_lumen.bin``, ``ini_ECM.bin``, ``data_int/lumen_step_<s>.dat``,
    ``data_int/ECM_step_<s>.dat``.  The ``.bin`` records take their second
    float from the slab, as the reference's writer does."""
    phiout, rhoout = slabs["phiout"], slabs["rhoout"]
    lumout, ecmout = slabs["lumout"], slabs["ECMout"]
    nf, nc, N = phiout.shape[0], phiout.shape[1], phiout.shape[2]
    ncur = nc if n_current is None else n_current
    os.makedirs(os.path.join(outdir, "data_int"), exist_ok=True)
    pl = N * N
    for it in range(nf):
        step = it * nsaving
        for n in range(ncur):
            start = it * pl * nc + n * pl
            write_bin(os.path.join(outdir, f"ini_cell_{n + 1}.bin"), phiout, start, pl)
            write_bin(os.path.join(outdir, f"ini_rho_{n + 1}.bin"), rhoout, start, pl)
            write_dat(os.path.join(outdir, "data_int", f"phi_cell_{n + 1}_step_{step}.dat"), phiout[it, n])
            write_dat(os.path.join(outdir, "data_int", f"rho_cell_{n + 1}_step_{step}.dat"), rhoout[it, n])
        write_bin(os.path.join(outdir, "ini_lumen.bin"), lumout, it * pl, pl)
        write_bin(os.path.join(outdir, "ini_ECM.bin"), ecmout, it * pl, pl)
        write_dat(os.path.join(outdir, "data_int", f"lumen_step_{step}.dat"), lumout[it])
        write_dat(os.path.join(outdir, "data_int", f"ECM_step_{step}.dat"), ecmout[it])


def read_state_files(indir: str, N: int, ncell: int) -> dict:
    """The state a run starts from (ini_phi_files, ini_rho_files,
    ini_lumen_files, ini_ECM_files; initialisation_functions.cpp:10-181):
    ``phi``, ``rho`` (ncell, N, N), ``lum``, ``ecm`` (N, N)."""
    pl = N * N
    phi = np.stack([read_bin(os.path.join(indir, f"ini_cell_{n + 1}.bin"), pl).reshape(N, N) for n in range(ncell)])
    rho = np.stack([read_bin(os.path.join(indir, f"ini_rho_{n + 1}.bin"), pl).reshape(N, N) for n in range(ncell)])
    lum = read_bin(os.path.join(indir, "ini_lumen.bin"), pl).reshape(N, N)
    ecm = read_bin(os.path.join(indir, "ini_ECM.bin"), pl).reshape(N, N)
    return {"phi": phi, "rho": rho, "lum": lum, "ecm": ecm}

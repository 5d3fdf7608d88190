"""Initial conditions and constants of the organoid workloads.

The reference reads its initial state from ``ini_*.bin`` files written by its
IC pipeline (functions/initialisation_functions.cpp:10-191,
initialisation_code/prepare_initial_files.cpp:56-118).  The workloads here are
synthetic and seeded; every array is built in float64 and handed to both the
GPU library (rounded to float32) and the CPU oracle.

Shapes:  phi, rho: (ncell, N, N);  lum, ECM: (N, N).

* ``phi_n = 0.5 * (1 - tanh((|x - c_n| - R_n) / eps))`` with ``x = (i, j)``, i
  the row index (the reference's "x", useful_functions.cpp:214).
* ``rho_n = 1 + 0.01 * U(-1, 1)`` where ``0.1 < phi_n < 0.9``, else 0.
* ``lum = 0`` and ``ECM = clip(1 - sum_n phi_n, 0, 1)`` (the outside-the-cyst
  branch of ini_ECM_lumen, initialisation_functions.cpp:84-114, without a
  lumen disc).

Random numbers come from SplitMix64 (Vigna), ``U[0,1) = (x >> 11) * 2**-53``,
seeded ``1000 + k`` for workload k, drawn in this order: centre jitter (2 per
cell, row then column), radii (1 per cell), then rho noise in row-major order
cell by cell over the band pixels only.

Hex packings: lattice spacing ``2 * mean(R)`` centred on the grid centre,
sites sorted by (distance, angle) and the first ``ncell`` kept; after jitter
the cells are ordered by angle around the grid centre (from -pi), because
the reference's "lateral" label uses index neighbours n-1, n+1 on a ring
(rho_functions.cpp:298-312).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from .params import OrgParams, grid_units, template_defaults

__all__ = ["SplitMix64", "Workload", "make_workload", "make_sweep", "WORKLOADS"]

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """SplitMix64 generator (Steele, Lea, Flood 2014; constants as published
    by Vigna).  Vectorised: the k-th output is mix(seed + k*gamma)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
        return lo + (hi - lo) * u


@dataclasses.dataclass
class Workload:
    name: str
    params: OrgParams
    phi: np.ndarray
    rho: np.ndarray
    lum: np.ndarray
    ecm: np.ndarray
    centres: np.ndarray       # (ncell, 2) float, (row, col)
    radii: np.ndarray         # (ncell,)
    s0_cell: np.ndarray | None
    steps: int

    @property
    def hfunc(self) -> np.ndarray:
        """The driver zeroes hfunc_cell before the loop (run_with_active_gels.cpp:74)."""
        return np.zeros_like(self.phi)


def _disc_fields(N, centres, radii, eps, rng):
    """phi_n = (1 - tanh((|x - c_n| - R_n) / eps)) / 2, rho_n = 1 + 0.01 U(-1, 1)
    on the band 0.1 < phi_n < 0.9 (row-major draw order), ECM = clip(1 - sum phi).
    Each cell is evaluated on the box where (|x - c| - R) / eps < 26: beyond
    it tanh is 1 in double precision and phi exactly 0, so the box-local
    result equals the full-plane one bit for bit (np.zeros pages that are
    never written cost no memory)."""
    nc = len(radii)
    phi = np.zeros((nc, N, N), np.float64)
    rho = np.zeros_like(phi)
    esum = np.zeros((N, N), np.float64)
    for n in range(nc):
        h = int(math.ceil(radii[n] + 26.0 * eps)) + 2
        i0, i1 = max(0, int(centres[n, 0]) - h), min(N, int(centres[n, 0]) + h + 1)
        j0, j1 = max(0, int(centres[n, 1]) - h), min(N, int(centres[n, 1]) + h + 1)
        ii, jj = np.meshgrid(np.arange(i0, i1, dtype=np.float64), np.arange(j0, j1, dtype=np.float64), indexing="ij")
        r = np.sqrt((ii - centres[n, 0]) ** 2 + (jj - centres[n, 1]) ** 2)
        box = 0.5 * (1.0 - np.tanh((r - radii[n]) / eps))
        phi[n, i0:i1, j0:j1] = box
        band = (box > 0.1) & (box < 0.9)
        cnt = int(band.sum())
        rb = np.zeros_like(box)
        rb[band] = 1.0 + 0.01 * rng.uniform(cnt, -1.0, 1.0)
        rho[n, i0:i1, j0:j1] = rb
        esum[i0:i1, j0:j1] += box
    lum = np.zeros((N, N), np.float64)
    ecm = np.clip(1.0 - esum, 0.0, 1.0)
    return phi, rho, lum, ecm


def hex_sites(ncell: int, spacing: float, centre: tuple[float, float]) -> np.ndarray:
    """The ``ncell`` hex-lattice sites nearest to ``centre`` (ties by angle)."""
    kmax = int(math.ceil(math.sqrt(ncell))) + 3
    pts = []
    for a in range(-kmax, kmax + 1):
        for b in range(-kmax, kmax + 1):
            x = spacing * (a + 0.5 * b)
            y = spacing * (b * math.sqrt(3.0) / 2.0)
            d = math.hypot(x, y)
            ang = math.atan2(y, x)
            pts.append((round(d, 9), ang, x, y))
    pts.sort()
    sel = np.array([[centre[0] + p[2], centre[1] + p[3]] for p in pts[:ncell]], np.float64)
    return sel


def _hex_workload(name, N, ncell, r_lo, r_hi, seed, steps, eps=2.0, disc=None, spacing=None):
    rng = SplitMix64(seed)
    c = (N / 2.0, N / 2.0)
    mean_r = 0.5 * (r_lo + r_hi)
    sites = hex_sites(ncell, spacing if spacing is not None else 2.0 * mean_r, c)
    jit = rng.uniform(2 * ncell, -2.0, 2.0).reshape(ncell, 2)
    centres = sites + jit
    if r_lo == r_hi:
        radii = np.full(ncell, float(r_lo))
    else:
        radii = rng.uniform(ncell, r_lo, r_hi)
    ang = np.arctan2(centres[:, 1] - c[1], centres[:, 0] - c[0])
    order = np.argsort(ang, kind="stable")
    centres, radii = centres[order], radii[order]
    phi, rho, lum, ecm = _disc_fields(N, centres, radii, eps, rng)
    p = grid_units(N, ncell, float(radii.mean()), eps=eps)
    s0_cell = None
    if r_lo != r_hi:
        p.use_s0_cell = 1
        s0_cell = math.pi * radii ** 2
    return Workload(name, p, phi, rho, lum, ecm, centres, radii, s0_cell, steps)


def label_constants(p: OrgParams) -> OrgParams:
    """Distinct apical / basal / lateral constants (each label's own deg, gen,
    a_ort, a_tan != a_ort, b and viscosity) around the template values, so
    that the turnover labels of create_turnover_rates_tensions
    (rho_functions.cpp:286-405) and the viscosity read at the mode's pixel
    (rho_functions.cpp:149-150, 170-186) change the results."""
    p.visc_ap, p.visc_bas, p.visc_lat = 0.5, 1.0, 2.0
    p.a_ort_ap, p.a_ort_bas, p.a_ort_lat = 80.0, 100.0, 120.0
    p.a_tan_ap, p.a_tan_bas, p.a_tan_lat = 80.0 * 1.25, 100.0 * 0.8, 120.0 * 1.5
    p.b_ap, p.b_bas, p.b_lat = 0.8, 1.0, 1.3
    p.gen_ap, p.gen_bas, p.gen_lat = 1.0, 1.5, 0.7
    p.deg_ap, p.deg_bas, p.deg_lat = 100.0, 120.0, 80.0
    return p


def _cyst_workload(name, N, ncell, R, seed, steps, eps=2.0, params=None):
    """Cells on a ring around a lumen, as the reference's initial state is
    built: ring radius rcyst = ncell * 0.9 * R / (pi dx) about (N/2, N/2) at
    angles 2 pi n / ncell (initialisation_functions.cpp:117-159, so ring
    neighbours are index neighbours), and where 1 - sum phi > 0 the lumen
    takes 1 - sum phi inside rcyst and the ECM outside
    (initialisation_functions.cpp:74-114).  The cell profile is this
    module's tanh disc (R and eps in pixels) with rho on the band.  params:
    the constants (default: grid units for R and eps)."""
    rng = SplitMix64(seed)
    rc = ncell * 0.9 * R / math.pi
    ang = 2.0 * math.pi * np.arange(ncell) / ncell
    centres = np.stack([N // 2 + rc * np.cos(ang), N // 2 + rc * np.sin(ang)], axis=1)
    radii = np.full(ncell, float(R))
    phi, rho, _, _ = _disc_fields(N, centres, radii, eps, rng)
    free = 1.0 - phi.sum(axis=0)
    ii, jj = np.mgrid[0:N, 0:N]
    d = np.sqrt((ii - N // 2) ** 2.0 + (jj - N // 2) ** 2.0)
    lum = np.where((free > 0.0) & (d < rc), free, 0.0)
    ecm = np.where((free > 0.0) & (d >= rc), free, 0.0)
    p = params if params is not None else grid_units(N, ncell, float(R), eps=eps)
    return Workload(name, p, phi, rho, lum, ecm, centres, radii, None, steps)


def _template_cyst(name, seed, steps, N=None):
    """The reference template's own configuration (constants_template.cpp
    with setup_simulation.py's defaults, params.template_defaults: 6 cells,
    dx = 0.01, R0 = R = 1, wi = 0.07, zone_size 200 so the FFT box is 398) in
    the analytic initial state of its IC program
    (prepare_initial_files.cpp:56-66), on the template's 1200^2 grid unless N
    is given:

    * phi_n = 1 / (1 + exp((dx |x - c_n| - R0) / wi)) with the centres on the
      ring of radius rcyst = N_ini 0.9 R / (pi dx) at angles 2 pi n / N_ini
      (initialisation_functions.cpp:117-159);
    * rho_n = rho0 exp(-1 / (4 h (1 - h) + tol) + 1 / (1 + tol)),
      h = phi_n^2 (3 - 2 phi_n), zero below rho0 exp(-1 / (2 tol) + 1 / (1 + tol)),
      rho0 = 0.001, tol = 0.01 (initialisation_functions.cpp:64-73;
      constants_template.cpp:110-111);
    * lum / ECM = 1 - sum phi where positive, inside / outside rcyst
      (initialisation_functions.cpp:82-114).

    Evaluated in float64 (the IC program's float rounding is not
    reproduced: the state is an input handed to both sides).  No random
    numbers: `seed` is unused."""
    del seed
    p = template_defaults()
    if N:
        p.N = int(N)
    N, nc = p.N, p.ncell
    rc = nc * 0.9 * p.R / (math.pi * p.dx)
    ang = 2.0 * math.pi * np.arange(nc) / nc
    centres = np.stack([N // 2 + rc * np.cos(ang), N // 2 + rc * np.sin(ang)], axis=1)
    ii = np.arange(N, dtype=np.float64)[:, None]
    jj = np.arange(N, dtype=np.float64)[None, :]
    phi = np.empty((nc, N, N))
    rho = np.empty((nc, N, N))
    rho0, tol = 0.001, 0.01
    floor = rho0 * math.exp(-1.0 / (2.0 * tol) + 1.0 / (1.0 + tol))
    for n in range(nc):
        r = np.sqrt((ii - centres[n, 0]) ** 2 + (jj - centres[n, 1]) ** 2)
        with np.errstate(over="ignore"):
            phi[n] = 1.0 / (1.0 + np.exp((p.dx * r - 1.0) / p.wi))
        h = phi[n] * phi[n] * (3.0 - 2.0 * phi[n])
        with np.errstate(divide="ignore"):
            v = rho0 * np.exp(-1.0 / (4.0 * h * (1.0 - h) + tol) + 1.0 / (1.0 + tol))
        rho[n] = np.where(v < floor, 0.0, v)
    free = 1.0 - phi.sum(axis=0)
    d = np.sqrt((ii - N // 2) ** 2 + (jj - N // 2) ** 2)
    lum = np.where((free > 0.0) & (d < rc), free, 0.0)
    ecm = np.where((free > 0.0) & (d >= rc), free, 0.0)
    radii = np.full(nc, p.R / p.dx)
    return Workload(name, p, phi, rho, lum, ecm, centres, radii, None, steps)


def make_workload(name: str, N: int | None = None, ncell: int | None = None, steps: int | None = None) -> Workload:
    """Build one of the named workloads.

    ``two_cells_256``     2 cells, R=30, eps=2, 256^2, centres 55 apart on row 128
    ``organoid16_512``    16 cells R=20 on a jittered hex packing, 512^2
    ``organoid64_1024``   64 cells, radii U(16,24), jittered hex packing, 1024^2
    ``organoid256_2048``  256 cells, radii U(14,26), jittered hex packing with
                          spacing 2 x 26 (no overlap at the start), 2048^2
    ``hex``               custom N / ncell hex packing with R=20 (tests)
    ``cyst6_256``         6 cells R=20 on a ring around a lumen, 256^2, with
                          distinct per-label constants (``label_constants``)
                          and lumen-cell coupling eta_lum_cells = 0.002
    ``two_cells_512_zone200``  two_cells_256's cells on 512^2 with the
                          template's zone_size 200 (FFT box 398)
    ``template6_1200``    the reference template's configuration: template
                          constants (6 cells, dx 0.01, zone_size 200, FFT box
                          398) in its IC program's analytic state (cells of
                          radius 100 pixels on the ring, rho from
                          ini_rho), 1200^2
    """
    if name == "two_cells_256":
        n_ = N or 256
        rng = SplitMix64(1000)
        centres = np.array([[n_ / 2.0, n_ / 2.0 - 27.5], [n_ / 2.0, n_ / 2.0 + 27.5]])
        radii = np.array([30.0, 30.0])
        phi, rho, lum, ecm = _disc_fields(n_, centres, radii, 2.0, rng)
        p = grid_units(n_, 2, 30.0, eps=2.0)
        return Workload(name, p, phi, rho, lum, ecm, centres, radii, None, steps or 1000)
    if name == "organoid16_512":
        return _hex_workload(name, N or 512, ncell or 16, 20.0, 20.0, 1001, steps or 5000)
    if name == "organoid64_1024":
        return _hex_workload(name, N or 1024, ncell or 64, 16.0, 24.0, 1003, steps or 10000)
    if name == "organoid256_2048":
        # spacing 2 * r_max: with the mean-radius spacing of the smaller
        # organoids, overlapping discs leave ECM pockets with near-flat
        # gradients, and the lumen/ECM reshaping loop (whose stop test is an
        # absolute threshold on the whole grid's change, so 256 cells need
        # hundreds of sweeps in the first step) diverges there -- the FP64
        # restatement of the reference does the same
        return _hex_workload(name, N or 2048, ncell or 256, 14.0, 26.0, 1004, steps or 20000, spacing=52.0)
    if name == "hex":
        return _hex_workload(name, N or 256, ncell or 4, 20.0, 20.0, 1099, steps or 10)
    if name == "two_cells_512_zone200":
        wl = make_workload("two_cells_256", N=512, steps=steps)
        wl.name = name
        wl.params.zone_size = 200
        return wl
    if name == "template6_1200":
        return _template_cyst(name, 1006, steps or 100, N=N)
    if name == "cyst6_256":
        wl = _cyst_workload(name, N or 256, ncell or 6, 20.0, 1005, steps or 100)
        label_constants(wl.params)
        wl.params.eta_lum_cells = 0.002
        return wl
    raise ValueError(f"unknown workload {name!r}")


def _fits(wl: Workload, margin: int = 6) -> bool:
    """Every cell's window plus a margin of 6 pixels (room to move) lies inside
    rows/cols 1..N-2 at the start."""
    zs, N = wl.params.zone_size, wl.params.N
    lo = np.floor(wl.centres).min() - zs - margin
    hi = np.ceil(wl.centres).max() + zs + margin
    return lo >= 1 and hi <= N - 2


SWEEP_SIZES = (256, 384, 512, 768, 1024)


def make_sweep(name: str = "sweep8", n_org: int = 8, steps: int = 2000) -> list:
    """An uneven batch of organoids (seed 1002): per organoid, in draw order,
    the grid side from {256, 384, 512, 768, 1024}, the cell count from
    4..64, the factors of a_ort, a_tan, b and friction, each log-uniform over
    +/- 0.5 decade around the template value, and a seed for its hex workload
    (radii U(12, 28), jitter, rho noise; see ``_hex_workload``).  A draw whose
    hex disc plus windows does not fit its grid is redrawn (cell count and
    seed) until it fits."""
    from .sweep import log_uniform
    if name != "sweep8":
        raise ValueError(f"unknown sweep {name!r}")
    rng = SplitMix64(1002)
    out = []
    for o in range(n_org):
        N = SWEEP_SIZES[int(rng.uniform(1)[0] * len(SWEEP_SIZES))]
        fa, ft, fb, ff = (log_uniform(float(u), 1.0) for u in rng.uniform(4))
        while True:
            ncell = 4 + int(rng.uniform(1)[0] * 61)
            seed = int(rng.next_u64(1)[0] & np.uint64(0x7FFFFFFF))
            wl = _hex_workload(f"{name}[{o}]", N, ncell, 12.0, 28.0, seed, steps)
            if _fits(wl):
                break
        p = wl.params
        p.a_ort_ap, p.a_ort_bas, p.a_ort_lat = p.a_ort_ap * fa, p.a_ort_bas * fa, p.a_ort_lat * fa
        p.a_tan_ap, p.a_tan_bas, p.a_tan_lat = p.a_tan_ap * ft, p.a_tan_bas * ft, p.a_tan_lat * ft
        p.b_ap, p.b_bas, p.b_lat = p.b_ap * fb, p.b_bas * fb, p.b_lat * fb
        p.friction = p.friction * ff
        out.append(wl)
    return out


WORKLOADS = ("two_cells_256", "organoid16_512", "organoid64_1024", "organoid256_2048")

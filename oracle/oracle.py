"""ctypes wrapper of the CPU oracle (organoid_oracle.c).  TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# field ids: ORG_F_* (include/organoid.h) and ORC_F_* (organoid_oracle.h)
F = dict(phi=0, rho=1, hfunc=2, lum=3, ecm=4, psi=5, hlum=6, hecm=7, lumn=8, ecmn=9,
         phin=10, rhon=11, ngp=12, pnd=13, jx=14, jy=15, temp=16, deg=17, gen=18,
         atan=19, aort=20, visc=21, b=22, pixx=23, pixy=24, piyy=25, nglum=26, jlx=27,
         jly=28, vx=29, vy=30)

_LIBS: dict[str, ctypes.CDLL] = {}


def load(precision: str = "f64") -> ctypes.CDLL:
    if precision in _LIBS:
        return _LIBS[precision]
    path = os.path.join(_HERE, f"liborganoid_oracle_{precision}.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: build it with `make -C oracle`")
    lib = ctypes.CDLL(path)
    vp, i32, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [vp, vp]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_field_size.restype = ctypes.c_int64
    lib.orc_field_size.argtypes = [vp, i32]
    for name in ("orc_set_field", "orc_get_field"):
        getattr(lib, name).argtypes = [vp, i32, vp]
    for name in ("orc_get_centres", "orc_set_centres", "orc_get_areas"):
        getattr(lib, name).argtypes = [vp, vp, vp]
    lib.orc_get_iterations.argtypes = [vp, vp, vp, vp]
    lib.orc_step.argtypes = [vp, i32]
    lib.orc_reshaping_lumen_or_ECM.argtypes = [vp, i32]
    for name in ("orc_init_like_driver", "orc_find_centres_cells", "orc_create_turnover_rates_tensions",
                 "orc_compute_hfunc_cells", "orc_compute_hfunc_lum_ecm", "orc_compute_normal_grad_phi",
                 "orc_compute_prod_norm_delta", "orc_compute_psi", "orc_sum_s", "orc_update_all_cortices",
                 "orc_update_lumen_and_ECM", "orc_compute_normal_to_phi", "orc_update_all_cells",
                 "orc_reshaping_cells", "orc_commit_step", "orc_last_error"):
        getattr(lib, name).argtypes = [vp]
    lib.orc_threads.restype = ctypes.c_int
    lib.orc_relax_failures.restype = ctypes.c_int64
    lib.orc_relax_failures.argtypes = [vp]
    lib.orc_fft2d.argtypes = [i32, vp, i32]
    lib.orc_fft1d.argtypes = [i32, vp, i32]
    lib.orc_bc_cells_planes.argtypes = [i32, i32, vp]
    lib.orc_velocity_modes.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp]
    _LIBS[precision] = lib
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """One organoid held by the CPU oracle, on full planes (float64 I/O)."""

    def __init__(self, params, s0_cell=None, precision: str = "f64"):
        self.lib = load(precision)
        self.params = params.copy()
        self.nc, self.N = params.ncell, params.N
        self._s0 = None if s0_cell is None else np.ascontiguousarray(s0_cell, np.float64)
        self.h = self.lib.orc_create(ctypes.byref(self.params), None if self._s0 is None else _ptr(self._s0))
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, name):
        N, M = self.N, self.params.M
        if name in ("vx", "vy"):                  # per-cell FFT boxes
            return (self.nc, M, M)
        n = self.lib.orc_field_size(self.h, F[name])
        if n == self.nc * N * N:
            return (self.nc, N, N)
        return (N, N)

    def set(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, np.float64)
        if a.size != self.lib.orc_field_size(self.h, F[name]):
            raise ValueError(f"{name}: size {a.size} != {self.lib.orc_field_size(self.h, F[name])}")
        self._check(self.lib.orc_set_field(self.h, F[name], _ptr(a)))

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self._shape(name), np.float64)
        self._check(self.lib.orc_get_field(self.h, F[name], _ptr(out)))
        return out

    def load_workload(self, wl) -> None:
        self.set("phi", wl.phi); self.set("rho", wl.rho)
        self.set("lum", wl.lum); self.set("ecm", wl.ecm)
        self.set("hfunc", np.zeros_like(wl.phi))
        self._check(self.lib.orc_init_like_driver(self.h))

    def centres(self):
        xc = np.zeros(self.nc, np.int32); yc = np.zeros(self.nc, np.int32)
        self.lib.orc_get_centres(self.h, _ptr(xc), _ptr(yc))
        return xc, yc

    def set_centres(self, xc, yc):
        xc = np.ascontiguousarray(xc, np.int32); yc = np.ascontiguousarray(yc, np.int32)
        self.lib.orc_set_centres(self.h, _ptr(xc), _ptr(yc))

    def areas(self):
        s = np.zeros(self.nc); f0 = np.zeros(self.nc)
        self.lib.orc_get_areas(self.h, _ptr(s), _ptr(f0))
        return s, f0

    def iterations(self):
        cells = np.zeros(self.nc, np.int32); lum = ctypes.c_int32(); ecm = ctypes.c_int32()
        self.lib.orc_get_iterations(self.h, _ptr(cells), ctypes.byref(lum), ctypes.byref(ecm))
        return cells, lum.value, ecm.value

    def call(self, fname: str, *args) -> int:
        return self._check(getattr(self.lib, "orc_" + fname)(self.h, *args))

    def step(self, it: int = 0) -> int:
        return self._check(self.lib.orc_step(self.h, it))

    def relax_failures(self) -> int:
        return int(self.lib.orc_relax_failures(self.h))

    def last_error(self) -> int:
        return self.lib.orc_last_error(self.h)

    @staticmethod
    def _check(rc):
        if rc:
            raise RuntimeError(f"oracle error code {rc}")
        return rc


def threads(precision: str = "f64") -> int:
    """OpenMP threads of the oracle's loops (cells, rows, pixels)."""
    return int(load(precision).orc_threads())


def fft2d(a: np.ndarray, sign: int, precision: str = "f64") -> np.ndarray:
    """Unnormalised 2D DFT of a complex square array via the oracle."""
    lib = load(precision)
    M = a.shape[0]
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float64)
    lib.orc_fft2d(M, _ptr(buf), sign)
    return buf[..., 0] + 1j * buf[..., 1]


def fft1d(a: np.ndarray, sign: int, precision: str = "f64") -> np.ndarray:
    lib = load(precision)
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float64)
    lib.orc_fft1d(a.shape[0], _ptr(buf), sign)
    return buf[..., 0] + 1j * buf[..., 1]


def bc_cells(planes: np.ndarray, precision: str = "f64") -> np.ndarray:
    lib = load(precision)
    a = np.ascontiguousarray(planes, np.float64).copy()
    nplanes = 1 if a.ndim == 2 else a.shape[0]
    lib.orc_bc_cells_planes(a.shape[-1], nplanes, _ptr(a))
    return a


def velocity_modes(params, pxx, pxy, pyy, eta, precision: str = "f64"):
    lib = load(precision)
    M = pxx.shape[0]
    c = lambda z: np.ascontiguousarray(np.stack([z.real, z.imag], -1), np.float64)
    a, b, d = c(pxx), c(pxy), c(pyy)
    e = np.ascontiguousarray(eta, np.float64)
    vx = np.zeros((M, M, 2)); vy = np.zeros((M, M, 2))
    lib.orc_velocity_modes(ctypes.byref(params), M, _ptr(a), _ptr(b), _ptr(d), _ptr(e), _ptr(vx), _ptr(vy))
    return vx[..., 0] + 1j * vx[..., 1], vy[..., 0] + 1j * vy[..., 1]

"""Python host side of the organoid library: ctypes over the C-ABI
(include/organoid.h, ``csrc/liborganoid_hip.so``).

``Organoid`` mirrors the reference's per-timestep interface: one method per
function the driver calls in its time loop (main_code/run_with_active_gels.cpp:
110-184), same names, same order, same meaning.  Functions whose outputs the
reference only consumes inside a cell's window (compute_normal_grad_phi,
compute_prod_norm_delta, create_turnover_rates_tensions, compute_normal_to_phi,
the boundary_conditions_* calls on those fields) are fused into the kernels
that consume them; their methods exist and do nothing, so driver-order code
reads like the reference.  ``step`` runs whole timesteps on the device
(batched form, optionally replayed from a hipGraph).

Errors: every C-ABI call returns a status; a non-zero status raises
``OrganoidError`` (the reference has no error channel and would read out of
range or divide by zero instead, see include/organoid.h).  There is no CPU
fallback: without the built library, or without a GPU, construction raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import OrgParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "csrc", "liborganoid_hip.so")

F_PHI, F_RHO, F_HFUNC, F_LUM, F_ECM, F_PSI, F_HLUM, F_HECM, F_LUMN, F_ECMN = range(10)
W_PHIN, W_RHON, W_VX, W_VY = range(4)
K_NAMES = ("centres", "prep", "psi", "lumecm", "lumreshape", "spectral", "cells", "commit", "step", "mid")
OPT_WRITE_VELOCITY, OPT_SURROGATE_JUDGE = 0, 1
JUDGE_SPECTRAL, JUDGE_FINITE_DIFF = 0, 1

_lib = None


class OrganoidError(RuntimeError):
    def __init__(self, code: int, what: str):
        self.code = code
        msg = _lib.org_error_string(code).decode() if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} (code {code})")


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the gfx950 library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: run __graft_entry__.build() (or make -C csrc)")
    try:
        # torch's wheel carries its own HIP runtime under the same soname;
        # loaded first, it is the one runtime the process uses (loaded
        # after this library, torch finds no GPU)
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    lib.org_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
    lib.org_create_f64.argtypes = [vp, vp, ctypes.POINTER(vp)]
    lib.org_destroy.argtypes = [vp]
    lib.org_get_precision.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    for n in ("org_set_field", "org_get_field", "org_get_window", "org_set_field_f64", "org_get_field_f64",
              "org_get_window_f64"):
        getattr(lib, n).argtypes = [vp, i32, vp]
    lib.org_get_centres.argtypes = [vp, vp, vp]
    lib.org_get_areas.argtypes = [vp, vp, vp]
    lib.org_get_iterations.argtypes = [vp, vp, vp, vp]
    for n in ("org_prepare", "org_find_centres_cells", "org_compute_hfunc_cells", "org_compute_hfunc_lum_ecm",
              "org_compute_psi", "org_sum_s", "org_update_all_cortices", "org_update_lumen_and_ECM",
              "org_update_all_cells", "org_reshaping_cells", "org_commit_step", "org_sync"):
        getattr(lib, n).argtypes = [vp]
    lib.org_reshaping_lumen_or_ECM.argtypes = [vp, i32]
    lib.org_reshaping_cells_from_phi.argtypes = [vp]
    lib.org_update_rho_to_shape.argtypes = [vp, ctypes.c_float, ctypes.c_float]
    lib.org_relax_initial_state.argtypes = [vp, ctypes.c_float, ctypes.c_float]
    lib.org_step.argtypes = [vp, i32, i32, i32]
    lib.org_graph_prepare.argtypes = [vp, i32]
    lib.org_profile_enable.argtypes = [vp, i32]
    lib.org_kernel_time.argtypes = [vp, i32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    lib.org_set_option.argtypes = [vp, i32, i32]
    lib.org_debug_fft2d.argtypes = [vp, i32]
    lib.org_debug_fft2d_generic.argtypes = [i32, vp, i32, i32, i32, i32]
    lib.org_get_fft_schedule.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.org_set_fft_schedule.argtypes = [vp, i32, i32]
    lib.org_batch_create.argtypes = [vp, i32, ctypes.POINTER(vp)]
    lib.org_batch_step.argtypes = [vp, i32, i32]
    lib.org_batch_sync.argtypes = [vp]
    lib.org_batch_destroy.argtypes = [vp]
    lib.org_get_tuning.argtypes = [vp, vp]
    lib.org_set_tuning.argtypes = [vp, vp]
    lib.org_surrogate_samples.argtypes = [vp, i32, i32, vp, vp, vp]
    lib.org_step_surrogate.argtypes = [vp, vp, vp, i32, i32, i32, ctypes.c_float, vp, vp]
    lib.org_get_relax_failures.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    lib.org_error_string.restype = ctypes.c_char_p
    lib.org_error_string.argtypes = [i32]
    lib.org_build_info.restype = ctypes.c_char_p
    _lib = lib
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _chk(rc: int, what: str) -> None:
    if rc:
        raise OrganoidError(rc, what)


class Tuning(ctypes.Structure):
    """org_tuning (include/organoid.h): launch shape of the batched step."""
    _fields_ = [("lum_blocks", ctypes.c_int32), ("roles", ctypes.c_int32), ("graph_steps", ctypes.c_int32),
                ("shared_bracket", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {"lum_blocks": self.lum_blocks, "roles": self.roles, "graph_steps": self.graph_steps,
                "shared_bracket": self.shared_bracket}


class Organoid:
    """One organoid resident on the GPU (library-owned device memory).

    precision "f32" (the reference's float build; zone_size 65 runs the LDS
    kernels, any other zone_size the full-plane path) or "f64" (the
    reference retyped to double, full-plane path).  Fields come back in the
    context's precision."""

    def __init__(self, params: OrgParams, s0_cell=None, precision: str = "f32"):
        if precision not in ("f32", "f64"):
            raise ValueError(f"precision {precision!r}: 'f32' or 'f64'")
        self.lib = load_library()
        self.params = params.copy()
        self.nc, self.N = params.ncell, params.N
        self.W, self.M = params.W, params.M
        self._s0 = None if s0_cell is None else np.ascontiguousarray(s0_cell, np.float64)
        h = ctypes.c_void_p()
        create = self.lib.org_create_f64 if precision == "f64" else self.lib.org_create
        _chk(create(ctypes.byref(self.params), None if self._s0 is None else _p(self._s0), ctypes.byref(h)),
             "org_create_f64" if precision == "f64" else "org_create")
        self.h = h
        rb = ctypes.c_int32()
        _chk(self.lib.org_get_precision(self.h, ctypes.byref(rb)), "org_get_precision")
        self.dtype = np.float64 if rb.value == 8 else np.float32
        # the per-function full-plane path (FP64, or an FFT box other than 128)
        self.generic = self.dtype == np.float64 or params.zone_size != 65

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.org_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state in / out -------------------------------------------------
    def _size(self, field: int) -> tuple:
        per_cell = field in (F_PHI, F_RHO, F_HFUNC)
        return (self.nc, self.N, self.N) if per_cell else (self.N, self.N)

    def set_field(self, field: int, arr) -> None:
        f64 = self.dtype == np.float64
        a = np.ascontiguousarray(arr, self.dtype)
        if a.shape != self._size(field):
            raise ValueError(f"field {field}: shape {a.shape} != {self._size(field)}")
        if f64:
            _chk(self.lib.org_set_field_f64(self.h, field, _p(a)), "org_set_field_f64")
        else:
            _chk(self.lib.org_set_field(self.h, field, _p(a)), "org_set_field")

    def get_field(self, field: int) -> np.ndarray:
        out = np.empty(self._size(field), self.dtype)
        if self.dtype == np.float64:
            _chk(self.lib.org_get_field_f64(self.h, field, _p(out)), "org_get_field_f64")
        else:
            _chk(self.lib.org_get_field(self.h, field, _p(out)), "org_get_field")
        return out

    def get_window(self, which: int) -> np.ndarray:
        side = self.W if which == W_PHIN else self.M
        out = np.empty((self.nc, side, side), self.dtype)
        if self.dtype == np.float64:
            _chk(self.lib.org_get_window_f64(self.h, which, _p(out)), "org_get_window_f64")
        else:
            _chk(self.lib.org_get_window(self.h, which, _p(out)), "org_get_window")
        return out

    def load_workload(self, wl) -> None:
        """Upload phi, rho, lum, ECM and a zero hfunc_cell, then seed the
        support boxes (the driver's copyin, run_with_active_gels.cpp:65-98)."""
        self.set_field(F_PHI, wl.phi)
        self.set_field(F_RHO, wl.rho)
        self.set_field(F_LUM, wl.lum)
        self.set_field(F_ECM, wl.ecm)
        self.set_field(F_HFUNC, np.zeros_like(wl.phi))
        self.prepare()

    def prepare(self) -> None:
        _chk(self.lib.org_prepare(self.h), "org_prepare")

    def centres(self):
        xc = np.zeros(self.nc, np.int32)
        yc = np.zeros(self.nc, np.int32)
        _chk(self.lib.org_get_centres(self.h, _p(xc), _p(yc)), "org_get_centres")
        return xc, yc

    def areas(self):
        s = np.zeros(self.nc)
        f0 = np.zeros(self.nc)
        _chk(self.lib.org_get_areas(self.h, _p(s), _p(f0)), "org_get_areas")
        return s, f0

    def iterations(self):
        cells = np.zeros(self.nc, np.int32)
        lum, ecm = ctypes.c_int32(), ctypes.c_int32()
        _chk(self.lib.org_get_iterations(self.h, _p(cells), ctypes.byref(lum), ctypes.byref(ecm)),
             "org_get_iterations")
        return cells, lum.value, ecm.value

    def relax_failures(self) -> int:
        """Calls in which a reshaping loop passed rimax (a warning, as the
        reference's "Relaxation failure!", cell_phase_functions.cpp:565, 862)."""
        n = ctypes.c_int64()
        _chk(self.lib.org_get_relax_failures(self.h, ctypes.byref(n)), "org_get_relax_failures")
        return n.value

    def fft_schedule(self):
        """(stages fused per LDS pass, lines per workgroup) of the full-plane
        path's transforms, as the autotuner at creation picked them; (0, 0)
        on the LDS path."""
        f, l = ctypes.c_int32(), ctypes.c_int32()
        _chk(self.lib.org_get_fft_schedule(self.h, ctypes.byref(f), ctypes.byref(l)), "org_get_fft_schedule")
        return f.value, l.value

    def set_fft_schedule(self, fuse: int, lines_per_wg: int) -> None:
        _chk(self.lib.org_set_fft_schedule(self.h, int(fuse), int(lines_per_wg)), "org_set_fft_schedule")

    def get_tuning(self) -> dict:
        t = Tuning()
        _chk(self.lib.org_get_tuning(self.h, ctypes.byref(t)), "org_get_tuning")
        return t.as_dict()

    def set_tuning(self, lum_blocks: int, roles: int, graph_steps: int, shared_bracket: int = 1) -> None:
        t = Tuning(lum_blocks, roles, graph_steps, shared_bracket)
        _chk(self.lib.org_set_tuning(self.h, ctypes.byref(t)), "org_set_tuning")

    def surrogate_samples(self, nsteps: int, it0: int = 0, residuals: bool = False):
        """Run nsteps exact batched steps and return, per step and cell, the
        (phi_n, rho_n) box crop at the step's start, the solve's (vx, vy) on
        the box ([nsteps*ncell, 2, M, M] CUDA tensors) and, if asked, the
        judge's residual of that exact velocity (org_surrogate_samples)."""
        import torch
        M = self.M
        x = torch.empty((nsteps * self.nc, 2, M, M), dtype=torch.float32, device="cuda")
        v = torch.empty_like(x)
        r = torch.empty(nsteps * self.nc, dtype=torch.float32, device="cuda") if residuals else None
        torch.cuda.synchronize()
        _chk(self.lib.org_surrogate_samples(self.h, it0, nsteps, ctypes.c_void_p(x.data_ptr()),
                                            ctypes.c_void_p(v.data_ptr()),
                                            ctypes.c_void_p(r.data_ptr()) if residuals else None),
             "org_surrogate_samples")
        self.sync()
        return (x, v, r) if residuals else (x, v)

    def step_surrogate(self, nsteps: int, net=None, v_given=None, k_exact: int = 10, threshold: float = 0.1,
                       it0: int = 0):
        """nsteps batched steps with the velocity from the FNO `net` (or the
        fixed CUDA tensor v_given [ncell, 2, M, M]), the exact solve every
        k_exact steps and wherever the judge's residual exceeds threshold
        (org_step_surrogate).  Returns (exact cells per step, residuals
        [nsteps, ncell])."""
        import torch
        exact = np.zeros(nsteps, np.int32)
        res = np.zeros((nsteps, self.nc), np.float32)
        torch.cuda.synchronize()
        _chk(self.lib.org_step_surrogate(self.h, net.h if net is not None else None,
                                         ctypes.c_void_p(v_given.data_ptr()) if v_given is not None else None,
                                         it0, nsteps, k_exact, float(threshold), _p(exact), _p(res)),
             "org_step_surrogate")
        return exact, res

    def set_option(self, option: int, value: int) -> None:
        _chk(self.lib.org_set_option(self.h, option, int(value)), "org_set_option")

    # ---- the driver's time-loop calls (run_with_active_gels.cpp:123-182) --
    def find_centres_cells(self):
        _chk(self.lib.org_find_centres_cells(self.h), "find_centres_cells")

    def create_turnover_rates_tensions(self):
        """Fused into update_all_cortices (the maps are only read there)."""

    def compute_hfunc_3d_tables_kernels_loop_n(self):
        _chk(self.lib.org_compute_hfunc_cells(self.h), "compute_hfunc_3d_tables_kernels_loop_n")

    def compute_all_hfunc_2d_tables_kernels(self):
        _chk(self.lib.org_compute_hfunc_lum_ecm(self.h), "compute_all_hfunc_2d_tables_kernels")

    def compute_normal_grad_phi(self):
        """Recomputed on the windows by the kernels that read it."""

    def boundary_conditions_cells(self):
        """Ghost rings of fields that are only read on the windows: no-op."""

    def compute_prod_norm_delta(self):
        """Fused into update_all_cortices."""

    def compute_psi_kernels(self):
        _chk(self.lib.org_compute_psi(self.h), "compute_psi_kernels")

    def sum_s_kernels(self):
        _chk(self.lib.org_sum_s(self.h), "sum_s_kernels")
        return self.areas()[0]

    def update_all_cortices(self):
        _chk(self.lib.org_update_all_cortices(self.h), "update_all_cortices")

    def update_lumen_and_ECM(self):
        _chk(self.lib.org_update_lumen_and_ECM(self.h), "update_lumen_and_ECM")

    def boundary_conditions_lumen_and_ECM(self):
        """Applied inside update_lumen_and_ECM."""

    def reshaping_lumen_or_ECM(self, which_ecm: int):
        _chk(self.lib.org_reshaping_lumen_or_ECM(self.h, int(which_ecm)), "reshaping_lumen_or_ECM")

    def compute_normal_to_phi(self):
        """Recomputed on the window by reshaping_cells."""

    def update_all_cells(self):
        _chk(self.lib.org_update_all_cells(self.h), "update_all_cells")

    def reshaping_cells(self):
        _chk(self.lib.org_reshaping_cells(self.h), "reshaping_cells")

    def copy_tables(self):
        """rho <- rhon, ECM <- ECMn, lum <- lumn (run_with_active_gels.cpp:175-182)."""
        _chk(self.lib.org_commit_step(self.h), "copy tables")

    # ---- the IC program's device part (prepare_initial_files.cpp:82-110) --
    def reshaping_cells_from_phi(self):
        """reshaping_cells with phin = phi (the IC program's copy, :87-90)."""
        _chk(self.lib.org_reshaping_cells_from_phi(self.h), "reshaping_cells (phin = phi)")

    def update_rho_to_shape(self, rho0: float = 0.001, tol: float = 0.01):
        """cell_phase_functions.cpp:42-62 on every plane."""
        _chk(self.lib.org_update_rho_to_shape(self.h, float(rho0), float(tol)), "update_rho_to_shape")

    def relax_initial_state(self, rho0: float = 0.001, tol: float = 0.01):
        """The whole IC relaxation: centres, cell reshaping from phi, rho to
        shape, lumen/ECM update (psi, h(lum), h(ECM) zero, as the program
        never computes them) and both reshaping loops; phi, lum and ECM end at
        their second-to-last iterates, as the program saves them."""
        _chk(self.lib.org_relax_initial_state(self.h, float(rho0), float(tol)), "relax_initial_state")

    def plain_step(self) -> None:
        """One timestep through the per-function entry points, in the
        driver's order (run_with_active_gels.cpp:123-182)."""
        self.find_centres_cells()
        self.create_turnover_rates_tensions()
        self.compute_hfunc_3d_tables_kernels_loop_n()
        self.compute_all_hfunc_2d_tables_kernels()
        self.compute_normal_grad_phi()
        self.boundary_conditions_cells()
        self.compute_prod_norm_delta()
        self.compute_psi_kernels()
        self.sum_s_kernels()
        self.update_all_cortices()
        self.update_lumen_and_ECM()
        self.boundary_conditions_lumen_and_ECM()
        self.reshaping_lumen_or_ECM(0)
        self.reshaping_lumen_or_ECM(1)
        self.compute_normal_to_phi()
        self.update_all_cells()
        self.reshaping_cells()
        self.copy_tables()

    # ---- batched form ---------------------------------------------------
    def step(self, nsteps: int = 1, graph: bool = False, it0: int = 0) -> None:
        _chk(self.lib.org_step(self.h, it0, nsteps, 1 if graph else 0), "org_step")

    def prepare_graph(self, nsteps: int) -> None:
        """Capture the hipGraphs a replay of nsteps steps uses, without running them."""
        _chk(self.lib.org_graph_prepare(self.h, int(nsteps)), "org_graph_prepare")

    def sync(self) -> None:
        _chk(self.lib.org_sync(self.h), "org_sync")

    def profile(self, on: bool) -> None:
        _chk(self.lib.org_profile_enable(self.h, 1 if on else 0), "org_profile_enable")

    def kernel_times(self) -> dict:
        out = {}
        for k, name in enumerate(K_NAMES):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            _chk(self.lib.org_kernel_time(self.h, k, ctypes.byref(ms), ctypes.byref(n)), "org_kernel_time")
            out[name] = (ms.value, n.value)
        return out


class Batch:
    """Organoids that advance together, one launch per stage for all of them
    (the counterpart of a setup_simulation.py sweep, whose runs the reference
    starts one by one, main_code/setup_simulation.py:66-98).  Each member
    evolves exactly as under ``Organoid.step``; the members stay usable
    (fields, centres, iteration counts) and must outlive the batch."""

    def __init__(self, members):
        self.lib = load_library()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h for m in self.members])
        h = ctypes.c_void_p()
        _chk(self.lib.org_batch_create(arr, len(self.members), ctypes.byref(h)), "org_batch_create")
        self.h = h

    @property
    def cells(self) -> int:
        return sum(m.nc for m in self.members)

    def step(self, nsteps: int = 1, graph: bool = True) -> None:
        _chk(self.lib.org_batch_step(self.h, int(nsteps), 1 if graph else 0), "org_batch_step")

    def sync(self) -> None:
        _chk(self.lib.org_batch_sync(self.h), "org_batch_sync")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.org_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def debug_fft2d_generic(a: np.ndarray, sign: int, f64: bool, fuse: int = 1, lines_per_wg: int = 1) -> np.ndarray:
    """M x M unnormalised 2D DFT through the full-plane path's transform
    (radix-2 or Bluestein lines in double; input and output stored as
    float64 when f64 else float32) with the given schedule."""
    lib = load_library()
    M = a.shape[0]
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float64)
    _chk(lib.org_debug_fft2d_generic(M, _p(buf), sign, 1 if f64 else 0, fuse, lines_per_wg),
         "org_debug_fft2d_generic")
    return buf[..., 0] + 1j * buf[..., 1]


def debug_fft2d(a: np.ndarray, sign: int) -> np.ndarray:
    """128 x 128 unnormalised 2D DFT through the library's LDS transform."""
    lib = load_library()
    buf = np.ascontiguousarray(np.stack([a.real, a.imag], -1), np.float32)
    _chk(lib.org_debug_fft2d(_p(buf), sign), "org_debug_fft2d")
    return buf[..., 0].astype(np.float64) + 1j * buf[..., 1]

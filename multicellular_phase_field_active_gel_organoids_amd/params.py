"""Simulation constants: the ctypes mirror of ``org_params`` (include/organoid.h).

The reference keeps its configuration as globals rendered from
``constants_template.cpp`` (functions/constants_template.cpp:1-130) by the
jinja sweep script (main_code/setup_simulation.py:15, 66-78).  Here one
``OrgParams`` value carries the same constants; ``template_defaults`` returns
the template values combined with the sweep defaults of setup_simulation.py:15.

``grid_units`` applies the one unit mapping used by every workload in this
package (grid spacing dx = 1, lengths in grid points).  The template is in
physical units (dx = 0.01, R0 = 1, wi = 0.07, constants_template.cpp:141-155,
233).  Restated quantities:

* ``wi = eps / sqrt(2)``: the reshaping flux (cell_phase_functions.cpp:511-547)
  is stationary on the logistic profile of width ``wi / sqrt(2)``; the
  workloads start from ``0.5 * (1 - tanh((r - R) / eps))``, a logistic of width
  ``eps / 2``, so this ``wi`` makes the initial profile the relaxed one.
* ``rdt`` keeps the reference's diffusion number of the reshaping step,
  ``wi**2 * rdt / dx**2 = 0.07**2 * 0.001 / 0.01**2 = 0.049``.
* ``rdev_threshold`` and ``rdev_ini`` (integers in the reference,
  constants_template.cpp:130-131) are scaled by ``(dx / 0.01)**2``: the lumen/ECM
  deviation ``sum |dlum| dx^2`` (cell_phase_functions.cpp:854) is an area while
  its threshold (:774) carries no length factor.
* ``R`` (reshaping tolerance radius, cell_phase_functions.cpp:473-475) is the
  mean cell radius in grid points and ``s0 = pi R^2``; per-cell target areas
  ``pi R_n^2`` are passed separately when radii differ.
* every dimensionless or rate constant keeps its template / sweep value.
"""
from __future__ import annotations

import ctypes
import math

__all__ = ["OrgParams", "template_defaults", "grid_units"]


class OrgParams(ctypes.Structure):
    """Field-for-field mirror of ``struct org_params`` in include/organoid.h."""

    _fields_ = [
        ("N", ctypes.c_int32),
        ("ncell", ctypes.c_int32),
        ("zone_size", ctypes.c_int32),
        ("rimax", ctypes.c_int32),
        ("rdev_threshold", ctypes.c_int32),
        ("rdev_ini", ctypes.c_int32),
        ("n_current", ctypes.c_int32),
        ("use_s0_cell", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("dx", ctypes.c_double),
        ("R", ctypes.c_double),
        ("s0", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("coupling_PF_AG", ctypes.c_double),
        ("tau", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("eta", ctypes.c_double),
        ("tau_lum", ctypes.c_double),
        ("beta_lum_cells", ctypes.c_double),
        ("eta_lum_cells", ctypes.c_double),
        ("xi_lum", ctypes.c_double),
        ("alpha_ECM", ctypes.c_double),
        ("tau_ECM", ctypes.c_double),
        ("beta_ECM_cells", ctypes.c_double),
        ("eta_ECM_cells", ctypes.c_double),
        ("xi_ECM", ctypes.c_double),
        ("beta_lum_ECM", ctypes.c_double),
        ("wi", ctypes.c_double),
        ("rtol", ctypes.c_double),
        ("rdt", ctypes.c_double),
        ("friction", ctypes.c_double),
        ("visc_ap", ctypes.c_double),
        ("visc_bas", ctypes.c_double),
        ("visc_lat", ctypes.c_double),
        ("a_ort_ap", ctypes.c_double),
        ("a_ort_bas", ctypes.c_double),
        ("a_ort_lat", ctypes.c_double),
        ("a_tan_ap", ctypes.c_double),
        ("a_tan_bas", ctypes.c_double),
        ("a_tan_lat", ctypes.c_double),
        ("b_ap", ctypes.c_double),
        ("b_bas", ctypes.c_double),
        ("b_lat", ctypes.c_double),
        ("gen_ap", ctypes.c_double),
        ("gen_bas", ctypes.c_double),
        ("gen_lat", ctypes.c_double),
        ("deg_ap", ctypes.c_double),
        ("deg_bas", ctypes.c_double),
        ("deg_lat", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

    def copy(self) -> "OrgParams":
        out = OrgParams()
        ctypes.pointer(out)[0] = self
        return out

    @property
    def M(self) -> int:
        """FFT box side, ``2*zone_size-2`` (run_with_active_gels.cpp:87)."""
        return 2 * self.zone_size - 2

    @property
    def W(self) -> int:
        """Stencil window side, ``2*zone_size`` (useful_functions.cpp:242-245)."""
        return 2 * self.zone_size


def template_defaults() -> OrgParams:
    """constants_template.cpp values with the sweep defaults of
    setup_simulation.py:15 substituted for the ``{{...}}`` placeholders."""
    p = OrgParams()
    p.N = 1200                   # constants_template.cpp:145
    p.ncell = 6                  # setup_simulation.py:15 ncells
    p.zone_size = 200            # :162
    p.rimax = 1000000            # :239
    p.rdev_threshold = 5         # :130 ("int rdev_threshold = 5.0")
    p.rdev_ini = 10              # :131
    p.n_current = p.ncell        # :160
    p.use_s0_cell = 0
    p.dt = 0.002                 # :142
    p.dx = 0.01                  # :141
    p.R = 1.0                    # :149-150
    p.s0 = 3.14159265359 * 1.0 * 1.0 + 0.01   # :151
    p.sigma = 0.2                # :155
    p.coupling_PF_AG = 1.0       # :166
    p.tau, p.alpha, p.beta = 1.0, 1.0, 1.0    # :208-210
    p.eta = 0.03                 # setup_simulation.py:15 eta
    p.tau_lum, p.beta_lum_cells = 1.0, 1.0    # :215-216
    p.eta_lum_cells = 0.0        # sweep eta_lum_cells
    p.xi_lum = 0.15              # sweep xi_lumen
    p.alpha_ECM = 0.002          # sweep alpha_ECM
    p.tau_ECM, p.beta_ECM_cells = 1.0, 1.0    # :223-224
    p.eta_ECM_cells = 0.001      # sweep eta_ecm_cells
    p.xi_ECM = 0.0               # :227
    p.beta_lum_ECM = 1.0         # :230
    p.wi = 0.07                  # :233
    p.rtol, p.rdt = 0.001, 0.001  # :241-242
    p.friction = 1.0             # sweep friction
    p.visc_ap = p.visc_bas = p.visc_lat = 1.0             # sweep viscosity_*
    p.a_ort_ap = p.a_ort_bas = p.a_ort_lat = 100.0         # sweep a_ort_*
    eps_ap = eps_bas = eps_lat = 1.0                       # sweep epsilon_*
    p.a_tan_ap, p.a_tan_bas, p.a_tan_lat = (p.a_ort_ap * eps_ap, p.a_ort_bas * eps_bas,
                                            p.a_ort_lat * eps_lat)   # :185-187
    p.b_ap = p.b_bas = p.b_lat = 1.0                       # sweep b_*
    p.gen_ap, p.deg_ap = 1.0, 100.0                        # sweep gen_ap, deg_ap
    p.gen_lat, p.deg_lat = 1.0, 100.0                      # :196-197
    p.gen_bas, p.deg_bas = p.gen_ap, p.deg_ap              # :198-199
    return p


def grid_units(N: int, ncell: int, R: float, eps: float = 2.0, zone_size: int = 65,
               dt: float = 1e-3) -> OrgParams:
    """Template constants restated in grid units (dx = 1); see module doc."""
    p = template_defaults()
    p.N = int(N)
    p.ncell = int(ncell)
    p.n_current = int(ncell)
    p.zone_size = int(zone_size)
    p.dx = 1.0
    p.dt = float(dt)
    p.R = float(R)
    p.s0 = math.pi * float(R) ** 2
    p.wi = float(eps) / math.sqrt(2.0)
    p.rdt = 0.049 / p.wi ** 2
    scale = (p.dx / 0.01) ** 2
    p.rdev_threshold = int(round(5 * scale))
    p.rdev_ini = 2 * p.rdev_threshold
    return p

"""Parameter sweeps as one batch of organoids.

The reference's sweep script (main_code/setup_simulation.py) takes a
dictionary of constants whose values are arrays, expands it into every
combination (create_constant_dictionnary, :29-50), renders one constants
file per combination from constants_template.cpp and prepares one run
directory each (:80-98); every run then steps alone.  Here the same
expansion gives one ``OrgParams`` per combination, and the organoids of a
sweep advance together in one ``host.Batch`` (one launch per stage).

Expansion order (setup_simulation.py:29-50): the key with the longest array
is chosen by ``np.argsort(lengths)[::-1][0]`` (numpy's default sort, as the
script calls it); for each of its values in order the key is fixed to that
scalar and the rest is expanded recursively, so that key varies slowest.
Keys whose value is not an array count as length 0.  A dictionary without
arrays expands to nothing, as in the script.
"""
from __future__ import annotations

import math

import numpy as np

from .params import OrgParams

__all__ = ["SWEEP_DEFAULTS", "create_constant_dictionnary", "constants_to_params", "sweep_params"]

# setup_simulation.py:15 (one value per key)
SWEEP_DEFAULTS = {
    "ncells": np.array([6]), "epsilon_bas": np.array([1.0]), "epsilon_lat": np.array([1.0]),
    "epsilon_ap": np.array([1.0]), "a_ort_bas": np.array([100.0]), "a_ort_lat": np.array([100.0]),
    "a_ort_ap": np.array([100.0]), "b_bas": np.array([1.0]), "b_lat": np.array([1.0]), "b_ap": np.array([1.0]),
    "gen_ap": np.array([1.0]), "deg_ap": np.array([100.0]), "viscosity_ap": np.array([1.0]),
    "viscosity_bas": np.array([1.0]), "viscosity_lat": np.array([1.0]), "friction": np.array([1.0]),
    "eta": np.array([0.03]), "xi_lumen": np.array([0.15]), "eta_lum_cells": np.array([0.0]),
    "alpha_ECM": np.array([0.002]), "eta_ecm_cells": np.array([0.001]),
}


def _lengths(dic: dict) -> list:
    return [len(v) if isinstance(v, np.ndarray) else 0 for v in dic.values()]


def create_constant_dictionnary(full: dict) -> list:
    """Every combination of the array-valued constants, in the script's order."""
    out = []
    dic = dict(full)
    keys = list(dic.keys())
    lens = _lengths(full)
    order = np.argsort(lens)[::-1]
    if lens[order[0]] > 0:
        key = keys[order[0]]
        for v in full[key]:
            dic[key] = v
            if sum(_lengths(dic)) == 0:
                out.append(dict(dic))
            else:
                out.extend(create_constant_dictionnary(dic))
    return out


def constants_to_params(dic: dict, base: OrgParams) -> OrgParams:
    """The template's use of the sweep placeholders (constants_template.cpp:34-101):
    a_tan = a_ort * epsilon per label, basal generation/degradation rates equal
    the apical ones, lateral ones keep the template values; the cell count
    comes from ``ncells``.  Everything else is taken from ``base``."""
    p = base.copy()
    f = lambda k: float(dic[k])
    p.ncell = int(dic["ncells"])
    p.n_current = p.ncell
    p.a_ort_ap, p.a_ort_bas, p.a_ort_lat = f("a_ort_ap"), f("a_ort_bas"), f("a_ort_lat")
    p.a_tan_ap = f("a_ort_ap") * f("epsilon_ap")
    p.a_tan_bas = f("a_ort_bas") * f("epsilon_bas")
    p.a_tan_lat = f("a_ort_lat") * f("epsilon_lat")
    p.b_ap, p.b_bas, p.b_lat = f("b_ap"), f("b_bas"), f("b_lat")
    p.gen_ap, p.deg_ap = f("gen_ap"), f("deg_ap")
    p.gen_bas, p.deg_bas = p.gen_ap, p.deg_ap
    p.visc_ap, p.visc_bas, p.visc_lat = f("viscosity_ap"), f("viscosity_bas"), f("viscosity_lat")
    p.friction = f("friction")
    p.eta = f("eta")
    p.xi_lum = f("xi_lumen")
    p.eta_lum_cells = f("eta_lum_cells")
    p.alpha_ECM = f("alpha_ECM")
    p.eta_ECM_cells = f("eta_ecm_cells")
    return p


def sweep_params(constants: dict, base: OrgParams) -> list:
    """One OrgParams per combination, in the script's order."""
    return [constants_to_params(d, base) for d in create_constant_dictionnary(constants)]


def log_uniform(u: float, centre: float, decades: float = 0.5) -> float:
    """centre * 10**(decades * (2u - 1)) for u in [0, 1)."""
    return centre * math.pow(10.0, decades * (2.0 * u - 1.0))

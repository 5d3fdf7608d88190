"""Pinning the oracle's conservative interface flux (reshape_update in
oracle/organoid_oracle.c), shared by the cell reshaping loop
(cell_phase_functions.cpp:498-549) and the lumen/ECM loop (:802-838).

The flux through a face is  n [wi^2 (n . grad u) - sqrt(2) wi u (1 - u)]  with
n the unit normal, so it vanishes on the logistic profile
u = 1 / (1 + exp(-sqrt(2) s / wi)), s the distance along n (width wi/sqrt(2)).
The discrete flux (face differences, the 0.25 cross-difference for the
tangential part of n . grad u, face averages) vanishes there to second order
in dx / wi.  Checked, on one sweep of each loop, for profiles laid along a
row and along the diagonal (where the cross-difference term carries half of
n . grad u):
  * the one-sweep change of the logistic of width wi/sqrt(2), relative to
    that of a profile of half or double width, is O((dx/wi)^2): halving
    dx/wi divides it by ~4, and at wi = 8 dx it is < 3 %;
  * profiles of width x0.5 and x2 are driven toward it (the sweep's change
    has a positive projection on the difference to the fixed point).
A wrong diffusion coefficient (wi instead of wi^2), a wrong compression
coefficient or a flipped cross-difference moves the fixed point and fails
the first check.
"""
import math

import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd.params import grid_units
from oracle.oracle import Oracle

N = 160
C = N // 2


def _params(wi):
    p = grid_units(N, 2, 20.0, zone_size=40)
    p.wi = float(wi)
    p.rdt = 0.049 / p.wi ** 2        # diffusion number of the template
    p.rtol = 1e6                      # stop after the first sweep
    p.rdev_threshold, p.rdev_ini = 1, 2
    return p


def _logistic(s, width):
    return 1.0 / (1.0 + np.exp(-s / width))


def _band(direction, wi, scale=1.0):
    """One interface of the profile of width scale * wi / sqrt(2) through the
    grid centre, rising along a row (direction 'row') or along the diagonal
    ('diag').  A single interface: two would interact where their tails
    overlap, which masks the second-order residual."""
    i, j = np.mgrid[0:N, 0:N].astype(np.float64)
    s = (i - C) if direction == "row" else ((i - C) + (j - C)) / math.sqrt(2.0)
    return _logistic(s, scale * wi / math.sqrt(2.0))


def _one_sweep_cells(u, wi):
    """One sweep of reshaping_cells on cell 0's window centred on the grid
    centre (normals from compute_normal_to_phi); returns phin - phi on the
    window."""
    o = Oracle(_params(wi))
    planes = np.stack([u, u])
    o.set("phi", planes)
    o.set("phin", planes)
    o.set_centres(np.array([C, C], np.int32), np.array([C, C], np.int32))
    o.call("compute_normal_grad_phi")
    o.call("compute_normal_to_phi")
    o.call("reshaping_cells")
    assert o.iterations()[0][0] == 1
    zs = o.params.zone_size
    win = (0, slice(C - zs, C + zs), slice(C - zs, C + zs))
    return (o.get("phin") - o.get("phi"))[win], u[win[1:]]


def _one_sweep_lumecm(u, wi, which_ecm):
    """One sweep of reshaping_lumen_or_ECM on the whole grid (normals from
    the new field over |grad| of the old one, equal here); the lumen takes
    u, the ECM 1 - u.  Returns the change of u on the interior away from the
    Dirichlet edges (one sweep disturbs only the pixels next to them)."""
    o = Oracle(_params(wi))
    f = 1.0 - u if which_ecm else u
    name, new = ("ecm", "ecmn") if which_ecm else ("lum", "lumn")
    o.set(name, f)
    o.set(new, f)
    o.call("reshaping_lumen_or_ECM", which_ecm)
    assert o.iterations()[2 if which_ecm else 1] == 1
    d = o.get(new) - f
    if which_ecm:
        d = -d
    inner = (slice(3, N - 3), slice(3, N - 3))
    return d[inner], u[inner]


LOOPS = {"cells": lambda u, wi: _one_sweep_cells(u, wi),
         "lumen": lambda u, wi: _one_sweep_lumecm(u, wi, 0),
         "ecm": lambda u, wi: _one_sweep_lumecm(u, wi, 1)}


@pytest.mark.parametrize("loop", sorted(LOOPS))
@pytest.mark.parametrize("direction", ["row", "diag"])
def test_logistic_profile_is_a_fixed_point(loop, direction):
    """The fixed point's one-sweep change relative to that of a profile of
    half or double width at the same wi: O((dx / wi)^2)."""
    sweep = LOOPS[loop]
    r = {}
    for wi in (4.0, 8.0):
        fixed = np.abs(sweep(_band(direction, wi), wi)[0]).max()
        wrong = min(np.abs(sweep(_band(direction, wi, sc), wi)[0]).max() for sc in (0.5, 2.0))
        r[wi] = fixed / wrong
    assert 2.5 < r[4.0] / r[8.0] < 6.5, r                     # second order in dx / wi
    assert r[8.0] < 0.03, r


@pytest.mark.parametrize("loop", sorted(LOOPS))
@pytest.mark.parametrize("direction", ["row", "diag"])
@pytest.mark.parametrize("scale", [0.5, 2.0])
def test_wrong_widths_are_driven_to_the_fixed_point(loop, direction, scale):
    wi = 6.0
    target = _band(direction, wi)
    d, u = LOOPS[loop](_band(direction, wi, scale), wi)
    zs = 40
    tgt = target[C - zs:C + zs, C - zs:C + zs] if loop == "cells" else target[3:N - 3, 3:N - 3]
    proj = float(np.sum(d * (tgt - u)))
    assert proj > 0.2 * float(np.linalg.norm(d) * np.linalg.norm(tgt - u)), proj

"""CPU restatement of the reference's initial-condition functions and file
formats -- TEST INFRASTRUCTURE (the checker of the drop-in's
initialisation functions and of the package's snapshot module).

Follows functions/initialisation_functions.cpp and the writers of
main_code/run_with_active_gels.cpp:191-269 and
initialisation_code/prepare_initial_files.cpp:116-200, with the reference's
expression types: float operands stay float32, a double literal promotes the
expression to float64, the stored result is float32.

Every function here is pinned by checks in tests/test_initial.py that do not
restate its formula: the logistic profile's half-value at R0 and its
symmetry, the ring of centres, the closed-form extremes of ini_rho, the
partition lum + ECM = max(0, 1 - sum phi), and byte-level layouts of the
files.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def cyst_radius(N_ini: int, R: float, dx: float) -> np.float32:
    """rcyst = N_ini*0.9*R/dx/M_PI, double arithmetic stored to float
    (initialisation_functions.cpp:76-77, 121-122)."""
    return f32(N_ini * 0.9 * float(f32(R)) / float(f32(dx)) / math.pi)


def ini_phi(N_s: int, N_ini: int, R: float, R0: float, dx: float, wi: float) -> np.ndarray:
    """Cells on a ring of radius rcyst around (N_s/2, N_s/2) at angles
    2*pi*n/N_ini, phi = 1/(1 + exp((dx*|x - c| - R0)/wi))
    (initialisation_functions.cpp:117-159).  One cell: centre (N_s/2, N_s/2),
    plane 0 only."""
    rc = cyst_radius(N_ini, R, dx)
    ii = np.arange(N_s, dtype=np.float32)[:, None]
    jj = np.arange(N_s, dtype=np.float32)[None, :]
    planes = N_ini if N_ini > 1 else 1
    out = np.zeros((planes, N_s, N_s), np.float32)
    for n in range(planes):
        if N_ini > 1:
            xc = f32((N_s // 2) + float(rc) * math.cos(n * 2.0 * math.pi / N_ini))
            yc = f32((N_s // 2) + float(rc) * math.sin(n * 2.0 * math.pi / N_ini))
        else:
            xc = yc = f32(N_s // 2)
        di = ii - xc
        dj = jj - yc
        r = np.sqrt(di * di + dj * dj)                     # float sqrt
        x = (f32(dx) * r - f32(R0)) / f32(wi)
        with np.errstate(over="ignore"):
            e = np.exp(x.astype(np.float64)).astype(np.float32)   # float exp, correctly rounded
        out[n] = (1.0 / (1.0 + e.astype(np.float64))).astype(np.float32)
    return out


def ini_rho(phi_n: np.ndarray, rho0: float, tol: float) -> np.ndarray:
    """rho = rho0*exp(-1/(4 h (1 - h) + tol) + 1/(1 + tol)) with h = phi^2(3 - 2 phi),
    set to 0 below rho0*exp(-1/(2 tol) + 1/(1 + tol))
    (initialisation_functions.cpp:64-73); double arithmetic."""
    p = phi_n.astype(np.float64)
    h = p * p * (3.0 - 2.0 * p)
    with np.errstate(divide="ignore"):
        v = float(f32(rho0)) * np.exp(-1.0 / (4.0 * h * (1.0 - h) + float(f32(tol))) + 1.0 / (1.0 + float(f32(tol))))
    rho = v.astype(np.float32)
    thr = float(f32(rho0)) * math.exp(-1.0 / (2.0 * float(f32(tol))) + 1.0 / (1.0 + float(f32(tol))))
    rho[rho.astype(np.float64) < thr] = 0.0
    return rho


def ini_ECM_lumen(phi: np.ndarray, lum: np.ndarray, ecm: np.ndarray, N_ini: int, R: float, dx: float):
    """Where 1 - sum_n phi_n > 0: lum = 1 - sum inside the cyst disc
    (d < rcyst around (N_s/2, N_s/2)), ECM = 1 - sum outside; elsewhere both
    0.  The other field keeps its value where one is written
    (initialisation_functions.cpp:82-114).  Returns new arrays."""
    N_s = phi.shape[-1]
    lum = lum.astype(np.float32, copy=True)
    ecm = ecm.astype(np.float32, copy=True)
    rc = cyst_radius(N_ini, R, dx)
    acc = np.zeros((N_s, N_s), np.float32)
    for n in range(N_ini):
        acc = acc + phi[n]                                 # float accumulation in cell order
    rest = 1.0 - acc.astype(np.float64)
    ii = np.arange(N_s)[:, None] - N_s // 2
    jj = np.arange(N_s)[None, :] - N_s // 2
    d = np.sqrt((ii * ii + jj * jj).astype(np.float64)).astype(np.float32)
    pos = rest > 0.0
    inside = d < rc
    lum[pos & inside] = rest[pos & inside].astype(np.float32)
    ecm[pos & ~inside] = rest[pos & ~inside].astype(np.float32)
    lum[~pos] = 0.0
    ecm[~pos] = 0.0
    return lum, ecm


def bin_bytes(slab: np.ndarray, start: int, count: int) -> bytes:
    """File bytes of the reference's binary writer for slab[start:start+count]:
    each element is written as sizeof(float*) = 8 bytes read from its address,
    i.e. the element and the next one of the slab
    (run_with_active_gels.cpp:197-201).  Past the end of the slab the
    reference reads unallocated memory; zeros stand in for it here."""
    s = np.ascontiguousarray(slab, np.float32).ravel()
    nxt = np.zeros(count, np.float32)
    avail = min(count, s.size - start - 1)
    if avail > 0:
        nxt[:avail] = s[start + 1:start + 1 + avail]
    rec = np.empty((count, 2), np.float32)
    rec[:, 0] = s[start:start + count]
    rec[:, 1] = nxt
    return rec.tobytes()


def bin_values(data: bytes, count: int) -> np.ndarray:
    """Values the reference's binary reader leaves in memory: element k is the
    first 4 bytes of record k (initialisation_functions.cpp:16-18; each
    8-byte read also writes element k+1, which the next read overwrites)."""
    a = np.frombuffer(data, np.float32, count=2 * count)
    return a[0::2].copy()


def dat_text(plane: np.ndarray, fmt: str = "%.15f") -> str:
    """Text snapshot: one value per line printed from the float promoted to
    double, a blank line after each row (run_with_active_gels.cpp:215-220)."""
    rows = []
    for row in np.asarray(plane, np.float32):
        rows.append("".join((fmt % float(v)) + "\n" for v in row) + "\n")
    return "".join(rows)


def update_rho_to_shape(phi: np.ndarray, rho: np.ndarray, rho0: float, tol: float) -> np.ndarray:
    """cell_phase_functions.cpp:42-62: where phi != 0, rho takes the ini_rho
    value of phi; then every value below rho0*exp(-1/(2 tol) + 1/(1 + tol))
    is set to 0 (also where phi == 0).  Double arithmetic."""
    out = np.array(rho, np.float32, copy=True)
    nz = phi != 0.0
    out[nz] = ini_rho(np.asarray(phi, np.float32)[nz], rho0, tol) if nz.any() else out[nz]
    # ini_rho already zeroes its values below the floor; the others:
    thr = float(f32(rho0)) * math.exp(-1.0 / (2.0 * float(f32(tol))) + 1.0 / (1.0 + float(f32(tol))))
    out[out.astype(np.float64) < thr] = 0.0
    return out


def relax_initial_state(o, rho0: float, tol: float) -> None:
    """The IC program's device part (prepare_initial_files.cpp:82-110) on an
    oracle.Oracle: centres, phin <- phi, |grad phi| and normals,
    reshaping_cells, update_rho_to_shape, then update_lumen_and_ECM with psi,
    h(lum), h(ECM) zero (the program never computes them) and both reshaping
    loops; no copy-back."""
    o.call("find_centres_cells")
    o.set("phin", o.get("phi"))
    o.call("compute_normal_grad_phi")
    o.call("compute_normal_to_phi")
    o.call("reshaping_cells")
    o.set("rho", update_rho_to_shape(o.get("phi"), o.get("rho"), rho0, tol))
    z = np.zeros((o.N, o.N))
    o.set("psi", z)
    o.set("hlum", z)
    o.set("hecm", z)
    o.call("update_lumen_and_ECM")
    o.call("reshaping_lumen_or_ECM", 0)
    o.call("reshaping_lumen_or_ECM", 1)

"""Initial-condition functions and state-file formats (host code, no GPU).

* The oracle's restatement (oracle/initial.py) is pinned by properties of the
  formulas it restates: the logistic profile is 1/2 at distance R0 and
  radially symmetric, the cells sit on the cyst ring, ini_rho peaks at rho0
  on the interface, vanishes in the bulk and is symmetric under
  phi -> 1 - phi, lum + ECM partitions 1 - sum phi, and the file layouts are
  checked byte by byte.
* The drop-in's C++ functions (csrc/dropin/organoid_initial.cpp, exercised by
  tests/dropin/ic_program) are checked against the oracle: phi within 2 ulp
  (float exp of the C library vs a correctly rounded one), rho, lum and ECM
  bit for bit, the reference-form .bin writer's bytes and the readers' values
  exactly, the .dat text byte for byte.
* The package's snapshot module (snapshots.py) writes the same bytes and
  reads the same values.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd import snapshots
from oracle import initial as oi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IC_PROGRAM = os.path.join(ROOT, "tests", "dropin", "ic_program")
# tests/dropin/constants_ic.cpp
IC = dict(N=96, nc=4, R=10.0, R0=6.0, dx=1.0, wi=1.0, rho0=0.001, tol=0.01, nsaving=7)


# ---- oracle pinning ---------------------------------------------------------
def test_ini_phi_profile_single_cell():
    N, R0, wi = 64, 10.0, 2.0
    phi = oi.ini_phi(N, 1, 10.0, R0, 1.0, wi)[0]
    c = N // 2
    assert phi[c, c + 10] == np.float32(0.5)            # exp(0) = 1 at dx*r = R0
    assert phi[c + 6, c + 8] == np.float32(0.5)         # r = 10 on a diagonal
    assert phi[c, c] == np.float32(1.0 / (1.0 + math.exp(-R0 / wi)))
    for a, b in ((3, 7), (12, 5), (0, 20)):
        v = phi[c + a, c + b]
        assert v == phi[c + b, c + a] == phi[c - a, c + b] == phi[c + a, c - b]
    row = phi[c, c:]
    assert np.all(np.diff(row.astype(np.float64)) <= 0)


def test_ini_phi_cells_on_the_cyst_ring():
    N, nc, R, R0 = 128, 6, 20.0, 8.0
    phi = oi.ini_phi(N, nc, R, R0, 1.0, 1.0)
    rc = nc * 0.9 * R / math.pi
    ii, jj = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    for n in range(nc):
        w = phi[n].astype(np.float64)
        ci, cj = (w * ii).sum() / w.sum(), (w * jj).sum() / w.sum()
        assert abs(ci - (N // 2 + rc * math.cos(2 * math.pi * n / nc))) < 0.02
        assert abs(cj - (N // 2 + rc * math.sin(2 * math.pi * n / nc))) < 0.02
        area = (phi[n] > 0.5).sum()
        assert abs(area - math.pi * R0 ** 2) < 2 * math.pi * R0   # a disc of radius R0


def test_ini_rho_closed_forms():
    rho0, tol = 0.001, 0.01
    p = np.array([0.0, 1.0, 0.5, 0.001, 0.999], np.float32)
    r = oi.ini_rho(p, rho0, tol)
    assert r[0] == 0 and r[1] == 0                       # bulk: exp(-1/tol + ...) below the floor
    assert r[2] == np.float32(rho0)                      # 4 h (1 - h) = 1 on the interface
    assert r[3] == 0 and r[4] == 0
    q = np.linspace(0.0, 1.0, 4097).astype(np.float32)
    rq = oi.ini_rho(q, rho0, tol)
    assert np.array_equal(rq, rq[::-1])                  # h(1 - phi) = 1 - h(phi)
    floor_v = rho0 * math.exp(-1.0 / (2 * tol) + 1.0 / (1 + tol))
    pos = rq[rq > 0].astype(np.float64)
    assert pos.min() >= floor_v * (1 - 1e-6) and pos.max() <= rho0 * (1 + 1e-7)
    half = rq[: 2049].astype(np.float64)
    assert np.all(np.diff(half) >= 0)                    # grows towards phi = 1/2


def test_ini_ECM_lumen_partition():
    N, nc, R = 96, 4, 10.0
    phi = oi.ini_phi(N, nc, R, 6.0, 1.0, 1.0)
    lum, ecm = oi.ini_ECM_lumen(phi, np.zeros((N, N)), np.zeros((N, N)), nc, R, 1.0)
    rest = 1.0 - phi.astype(np.float32).sum(axis=0, dtype=np.float32).astype(np.float64)
    want = np.where(rest > 0, rest, 0.0).astype(np.float32)
    assert np.array_equal(lum + ecm, want)
    assert np.all((lum == 0) | (ecm == 0))
    rc = nc * 0.9 * R / math.pi
    ii, jj = np.meshgrid(np.arange(N) - N // 2, np.arange(N) - N // 2, indexing="ij")
    d = np.sqrt(ii ** 2 + jj ** 2)
    assert np.all(lum[d >= rc + 1e-6] == 0) and np.all(ecm[d < rc - 1e-6] == 0)
    assert (lum > 0).sum() > 0 and (ecm > 0).sum() > 0
    # a field that is not written keeps its value
    l2, e2 = oi.ini_ECM_lumen(phi, np.zeros((N, N)), np.full((N, N), 7.0), nc, R, 1.0)
    keep = (rest > 0) & (d < rc)
    assert np.all(e2[keep] == 7.0)


def test_update_rho_to_shape_closed_forms():
    rho0, tol = 0.001, 0.01
    floor_v = rho0 * math.exp(-1.0 / (2 * tol) + 1.0 / (1 + tol))
    phi = np.array([0.0, 0.0, 0.0, 0.5, 1.0, 0.3], np.float32)
    rho = np.array([0.5, np.float32(floor_v) * 0.5, 7.0, 3.0, 3.0, 3.0], np.float32)
    out = oi.update_rho_to_shape(phi, rho, rho0, tol)
    assert out[0] == np.float32(0.5) and out[2] == np.float32(7.0)   # phi == 0: kept above the floor
    assert out[1] == 0.0                                             # ... and zeroed below it
    assert out[3] == np.float32(rho0) and out[4] == 0.0              # interface peak, bulk
    assert out[5] == oi.ini_rho(phi[5:6], rho0, tol)[0]


def test_bin_layout_and_dat_text():
    slab = np.arange(10, dtype=np.float32) + 0.25
    raw = oi.bin_bytes(slab, 6, 4)
    rec = np.frombuffer(raw, np.float32).reshape(4, 2)
    assert len(raw) == 32
    assert list(rec[:, 0]) == [6.25, 7.25, 8.25, 9.25]
    assert list(rec[:, 1]) == [7.25, 8.25, 9.25, 0.0]
    assert np.array_equal(oi.bin_values(raw, 4), slab[6:10])
    txt = oi.dat_text(np.array([[0.5, 1.0 / 3.0], [0.0, 1.0]], np.float32))
    assert txt == "0.500000000000000\n0.333333343267441\n\n0.000000000000000\n1.000000000000000\n\n"


# ---- the package's snapshot module --------------------------------------------
def test_snapshots_match_the_oracle(tmp_path):
    rng = np.random.default_rng(3)
    slab = rng.random((2, 3, 8, 8)).astype(np.float32)
    p = str(tmp_path / "x.bin")
    for start in (0, 64, 5 * 64):
        snapshots.write_bin(p, slab, start, 64)
        data = open(p, "rb").read()
        assert data == oi.bin_bytes(slab, start, 64)
        assert np.array_equal(snapshots.read_bin(p, 64), slab.ravel()[start:start + 64])
    with pytest.raises(ValueError):
        snapshots.read_bin(p, 65)
    q = str(tmp_path / "x.dat")
    plane = (rng.random((8, 8)) * 0.98 + 0.01).astype(np.float32)
    snapshots.write_dat(q, plane)
    assert open(q).read() == oi.dat_text(plane)
    back = snapshots.read_dat(q, 8)
    assert np.max(np.abs(back - plane)) <= 1e-15 + np.spacing(plane).max()
    assert snapshots.frame_count(11, 10) == 2 and snapshots.frame_count(10, 10) == 1


# ---- drop-in C++ functions ---------------------------------------------------
@pytest.fixture(scope="module")
def ic_dir(tmp_path_factory):
    assert os.path.exists(IC_PROGRAM), "build() builds tests/dropin/ic_program"
    d = tmp_path_factory.mktemp("ic")
    subprocess.run([IC_PROGRAM, str(d)], check=True, timeout=60)
    return str(d)


def _f32(d, name, shape):
    return np.fromfile(os.path.join(d, name), np.float32).reshape(shape)


def test_dropin_initialisers_match_the_oracle(ic_dir):
    N, nc = IC["N"], IC["nc"]
    phi = _f32(ic_dir, "ic_phi.f32", (nc, N, N))
    want = oi.ini_phi(N, nc, IC["R"], IC["R0"], IC["dx"], IC["wi"])
    assert np.all(np.abs(phi - want) <= 2 * np.spacing(want))
    rho = _f32(ic_dir, "ic_rho.f32", (nc, N, N))
    assert np.array_equal(rho, np.stack([oi.ini_rho(phi[n], IC["rho0"], IC["tol"]) for n in range(nc)]))
    lum, ecm = oi.ini_ECM_lumen(phi, np.zeros((N, N)), np.zeros((N, N)), nc, IC["R"], IC["dx"])
    assert np.array_equal(_f32(ic_dir, "ic_lum.f32", (N, N)), lum)
    assert np.array_equal(_f32(ic_dir, "ic_ecm.f32", (N, N)), ecm)
    assert (lum > 0).any() and (ecm > 0).any() and (rho > 0).any()


def _slabs(ic_dir):
    N, nc = IC["N"], IC["nc"]
    phi = _f32(ic_dir, "ic_phi.f32", (nc, N, N))
    rho = _f32(ic_dir, "ic_rho.f32", (nc, N, N))
    lum = _f32(ic_dir, "ic_lum.f32", (N, N))
    ecm = _f32(ic_dir, "ic_ecm.f32", (N, N))
    half = np.float32(0.5)
    return {"phiout": np.stack([phi, phi * half]), "rhoout": np.stack([rho, rho * half]),
            "lumout": np.stack([lum, lum * half]), "ECMout": np.stack([ecm, ecm * half])}


def test_dropin_files_and_readers(ic_dir, tmp_path):
    N, nc = IC["N"], IC["nc"]
    slabs = _slabs(ic_dir)
    pl = N * N
    # reference-form writer: the last frame's bytes, second halves from the slab
    # (zero past its end: ic_program allocates one extra zero float)
    for n in range(nc):
        raw = open(os.path.join(ic_dir, f"ini_cell_{n + 1}.bin"), "rb").read()
        assert raw == oi.bin_bytes(slabs["phiout"], nc * pl + n * pl, pl)
    raw = open(os.path.join(ic_dir, "ini_ECM.bin"), "rb").read()
    assert raw == oi.bin_bytes(slabs["ECMout"], pl, pl)
    # the drop-in's readers give the last saved frame
    assert np.array_equal(_f32(ic_dir, "rb_phi.f32", (nc, N, N)), slabs["phiout"][1])
    assert np.array_equal(_f32(ic_dir, "rb_rho.f32", (nc, N, N)), slabs["rhoout"][1])
    assert np.array_equal(_f32(ic_dir, "rb_lum.f32", (N, N)), slabs["lumout"][1])
    assert np.array_equal(_f32(ic_dir, "rb_ecm.f32", (N, N)), slabs["ECMout"][1])
    # .dat text as the oracle prints it
    dat = open(os.path.join(ic_dir, "data_int", f"rho_cell_2_step_{IC['nsaving']}.dat")).read()
    assert dat == oi.dat_text(slabs["rhoout"][1, 1])
    # the package writes the same files and reads the same state
    out = str(tmp_path / "run")
    snapshots.write_run_outputs(out, slabs, IC["nsaving"])
    names = sorted(os.listdir(os.path.join(ic_dir, "data_int")))
    assert names == sorted(os.listdir(os.path.join(out, "data_int")))
    for name in names:
        assert open(os.path.join(ic_dir, "data_int", name)).read() == open(os.path.join(out, "data_int", name)).read()
    for name in [f"ini_cell_{n + 1}.bin" for n in range(nc)] + [f"ini_rho_{n + 1}.bin" for n in range(nc)]:
        a = open(os.path.join(ic_dir, name), "rb").read()
        b = open(os.path.join(out, name), "rb").read()
        # the package writes 0 past the slab's end, ic_program its extra zero float
        assert a == b, name
    st = snapshots.read_state_files(out, N, nc)
    assert np.array_equal(st["phi"], slabs["phiout"][1]) and np.array_equal(st["ecm"], slabs["ECMout"][1])

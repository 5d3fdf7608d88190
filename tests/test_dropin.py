"""The drop-in boundary: the C-ABI library exports every entry point its
header declares; the C++ drop-in exports the reference's functions with
their C++ signatures and links a program written in the reference driver's
call order; on the GPU that program reproduces the batched form bit for bit.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from multicellular_phase_field_active_gel_organoids_amd.params import grid_units

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "multicellular_phase_field_active_gel_organoids_amd")
LIB = os.path.join(PKG, "csrc", "liborganoid_hip.so")
DROPIN = os.path.join(PKG, "csrc", "dropin", "liborganoid_dropin.so")
DROPIN64 = os.path.join(PKG, "csrc", "dropin", "liborganoid_dropin_f64.so")
DRIVER = os.path.join(ROOT, "tests", "dropin", "driver_two_cells")
DRIVER64 = os.path.join(ROOT, "tests", "dropin", "driver_two_cells_f64")


def _exports(path, demangle=False):
    cmd = ["nm", "-D", "--defined-only"] + (["-C"] if demangle else []) + [path]
    out = subprocess.run(cmd, check=True, capture_output=True, text=True).stdout
    return out


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "organoid.h")).read()
    declared = set(re.findall(r"^(?:int|const char \*)\s*\*?\s*(org_\w+)\(", hdr, re.M))
    assert len(declared) >= 25
    exported = set(re.findall(r" T (org_\w+)$", _exports(LIB), re.M))
    assert declared <= exported, declared - exported


def test_fno_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "organoid_fno.h")).read()
    declared = set(re.findall(r"^(?:int|size_t)\s+(fno_\w+)\(", hdr, re.M))
    assert len(declared) == 15
    exported = set(re.findall(r" T (fno_\w+)$", _exports(LIB), re.M))
    assert declared <= exported, declared - exported


@pytest.mark.parametrize("lib,real,cplx", [(DROPIN, "float", "org_complex32"), (DROPIN64, "double", "org_complex64")],
                         ids=["f32", "f64"])
def test_dropin_exports_reference_signatures(lib, real, cplx):
    """The reference's functions with their C++ signatures: float for the
    reference as written, double (and a {double x, y} FFT buffer tag) for the
    reference retyped to double."""
    out = _exports(lib, demangle=True)
    r, c = real, cplx
    for sig in [
        f"update_all_cortices({r}*, {r}*, {r}*, int*, int*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, "
        f"{r}*, {r}*, int, int, {c}*, {c}*, {c}*, {c}*, {c}*, {c}*, {c}*, {c}*, {c}*, {c}*, {c}*)",
        f"update_all_cells(int, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, int*, int*, {r}*, {r}*)",
        f"reshaping_cells(int, {r}*, {r}*, {r}*, {r}*, {r}*, int*, int*)",
        f"reshaping_lumen_or_ECM(int, {r}*, {r}*, {r})",
        f"update_lumen_and_ECM({r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*)",
        f"create_turnover_rates_tensions({r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, int*, int*)",
        f"compute_prod_norm_delta({r}*, {r}*, {r}*, int*, int*)",
        f"find_centres_cells({r}*, int*, int*)", f"sum_s_kernels({r}*, {r}*, {r}*)",
        f"saving_out_functions(int, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*, {r}*)",
        f"copy_cell_tables_1_by_1(int, {r}*, {r}*)", f"copy_tables_ECM_or_lumen({r}*, {r}*)",
        f"update_rho_to_shape({r}*, {r}*, {r}, {r})", f"ini_rho({r}*, {r}*, {r}, {r}, int)",
        "malloc_all()", "ini_fft_all()", "cufftPlan2d(int*, int, int, int)",
    ]:
        assert sig in out, sig
    for glob in ("phi", "rhon", "hfunc_cell", "temp_fft", "fft_rho_vy", "planf", "xcenters"):
        assert re.search(rf" [BD] {glob}$", out, re.M), glob


@pytest.mark.parametrize("name", ["constants_two_cells.cpp", "constants_ic_relax.cpp"])
def test_dropin_constants_match_workload_params(name):
    p = grid_units(256, 2, 30.0)
    src = open(os.path.join(ROOT, "tests", "dropin", name)).read()
    vals = dict(re.findall(r"^(?:const )?(?:int|float|org_real_t) (\w+) = ([^;]+);", src, re.M))
    assert int(vals["N_s"]) == p.N and int(vals["Ntot"]) == p.ncell and int(vals["zone_size"]) == p.zone_size
    for name, attr in (("dt", "dt"), ("dx", "dx"), ("wi", "wi"), ("rdt", "rdt"), ("s0", "s0"), ("R", "R"),
                       ("eta", "eta"), ("alpha_ECM", "alpha_ECM"), ("friction", "friction")):
        assert np.float32(float(vals[name])) == np.float32(getattr(p, attr)), name
    assert int(vals["rdev_threshold"]) == p.rdev_threshold and int(vals["rdev_ini"]) == p.rdev_ini


@pytest.mark.parametrize("prog,lib", [(DRIVER, "liborganoid_dropin.so"), (DRIVER64, "liborganoid_dropin_f64.so")],
                         ids=["f32", "f64"])
def test_driver_program_links(prog, lib):
    assert os.path.exists(prog), "build() builds the programs under tests/dropin"
    out = subprocess.run(["ldd", prog], capture_output=True, text=True).stdout
    assert lib in out and "not found" not in out


@pytest.mark.gpu
def test_driver_order_program_matches_batched_step(tmp_path):
    from multicellular_phase_field_active_gel_organoids_amd import host
    from multicellular_phase_field_active_gel_organoids_amd.configs import make_workload

    wl = make_workload("two_cells_256")
    for name, arr in (("phi", wl.phi), ("rho", wl.rho), ("lum", wl.lum), ("ecm", wl.ecm)):
        arr.astype(np.float32).tofile(tmp_path / f"{name}.f32")
    print("case drop-in driver two_cells_256, 3 steps", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    g = host.Organoid(wl.params)
    g.load_workload(wl)
    g.step(1)
    # the reference's diagnostics at step 0 (ndisplay = 5): window bounds and
    # reshaping counts per cell (cell_phase_functions.cpp:463-471, 575-578)
    xc, yc = g.centres()
    cells, _, _ = g.iterations()
    zs = wl.params.zone_size
    printed = [int(v) for v in re.findall(r"count_ri = (\d+)", r.stderr)]
    assert printed == [int(c) for c in cells], (printed, cells)
    for n in range(wl.params.ncell):
        assert f"xcenters[{n}]-{zs} = {xc[n] - zs}" in r.stderr and f"ycenters[{n}]+{zs} = {yc[n] + zs}" in r.stderr
    g.step(2)
    nc, N = wl.params.ncell, wl.params.N
    for name, f, shape in (("phi", host.F_PHI, (nc, N, N)), ("rho", host.F_RHO, (nc, N, N)),
                           ("lum", host.F_LUM, (N, N)), ("ecm", host.F_ECM, (N, N))):
        got = np.fromfile(tmp_path / f"out_{name}.f32", np.float32).reshape(shape)
        assert np.array_equal(got, g.get_field(f)), name


# org_params field <- the constants.h global the drop-in reads it from
# (organoid_dropin.cpp params_from_globals)
GLOBALS = dict(N="N_s", ncell="Ntot", zone_size="zone_size", rimax="rimax", rdev_threshold="rdev_threshold",
               rdev_ini="rdev_ini", n_current="N_current", dt="dt", dx="dx", R="R", s0="s0", sigma="sigma",
               coupling_PF_AG="coupling_PF_AG", tau="tau", alpha="alpha", beta="beta", eta="eta",
               tau_lum="tau_lum", beta_lum_cells="beta_lum_cells", eta_lum_cells="eta_lum_cells", xi_lum="xi_lum",
               alpha_ECM="alpha_ECM", tau_ECM="tau_ECM", beta_ECM_cells="beta_ECM_cells",
               eta_ECM_cells="eta_ECM_cells", xi_ECM="xi_ECM", beta_lum_ECM="beta_lum_ECM", wi="wi", rtol="rtol",
               rdt="rdt", friction="friction", visc_ap="viscosity_ap", visc_bas="viscosity_bas",
               visc_lat="viscosity_lat", a_ort_ap="a_ort_apical", a_ort_bas="a_ort_basal", a_ort_lat="a_ort_lateral",
               a_tan_ap="a_tan_apical", a_tan_bas="a_tan_basal", a_tan_lat="a_tan_lateral", b_ap="b_ap",
               b_bas="b_bas", b_lat="b_lat", gen_ap="matgenrate_apical", gen_bas="matgenrate_basal",
               gen_lat="matgenrate_lateral", deg_ap="degrate_apical", deg_bas="degrate_basal",
               deg_lat="degrate_lateral")


def params_from_constants(name, base):
    """The org_params a program linked with tests/dropin/<name> hands the
    library: the file's literals read as doubles (the FP64 form keeps them
    unrounded, so they must come from the file, not from the workload)."""
    src = open(os.path.join(ROOT, "tests", "dropin", name)).read()
    vals = dict(re.findall(r"^(?:const )?(?:int|float|org_real_t) (\w+) = ([^;]+);", src, re.M))
    p = base.copy()
    for field, glob in GLOBALS.items():
        v = float(vals[glob])
        setattr(p, field, int(v) if isinstance(getattr(p, field), int) else v)
    return p


@pytest.mark.gpu
def test_f64_driver_order_program_matches_f64_context_and_oracle(tmp_path):
    """The driver's call order through liborganoid_dropin_f64.so (the
    reference retyped to double) gives, bit for bit, what an FP64 context
    stepped by org_step gives from the same constants, and stays within
    1e-12 of the FP64 oracle."""
    from multicellular_phase_field_active_gel_organoids_amd import host
    from multicellular_phase_field_active_gel_organoids_amd.configs import make_workload
    from oracle.oracle import Oracle

    wl = make_workload("two_cells_256")
    wl.params = params_from_constants("constants_two_cells.cpp", wl.params)
    for name, arr in (("phi", wl.phi), ("rho", wl.rho), ("lum", wl.lum), ("ecm", wl.ecm)):
        np.asarray(arr, np.float64).tofile(tmp_path / f"{name}.f64")
    print("case FP64 drop-in driver two_cells_256, 3 steps", flush=True)
    r = subprocess.run([DRIVER64, str(tmp_path), "3"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    g = host.Organoid(wl.params, precision="f64")
    g.load_workload(wl)
    g.step(3)
    o = Oracle(wl.params)
    o.load_workload(wl)
    for it in range(3):
        o.step(it)
    nc, N = wl.params.ncell, wl.params.N
    for name, f, shape in (("phi", host.F_PHI, (nc, N, N)), ("rho", host.F_RHO, (nc, N, N)),
                           ("lum", host.F_LUM, (N, N)), ("ecm", host.F_ECM, (N, N))):
        got = np.fromfile(tmp_path / f"out_{name}.f64", np.float64).reshape(shape)
        assert np.array_equal(got, g.get_field(f)), name
        ref = o.get(name)
        r_o = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)
        print(f"  {name}: vs FP64 oracle {r_o:.2e}", flush=True)
        assert r_o <= 1e-12, (name, r_o)
    o.close()


@pytest.mark.gpu
def test_ic_program_order_matches_relax_initial_state(tmp_path):
    """The IC program's call order through the drop-in (phin <- phi marks
    the reshaping that starts from phi; update_rho_to_shape; lumen/ECM with
    psi, h zero; no copy-back) gives the saved frame org_relax_initial_state
    gives, bit for bit, from the same initial arrays."""
    from multicellular_phase_field_active_gel_organoids_amd import host
    prog = os.path.join(ROOT, "tests", "dropin", "ic_relax_program")
    print("case drop-in IC program, 256^2, 2 cells", flush=True)
    r = subprocess.run([prog, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    p = grid_units(256, 2, 30.0)
    nc, N = p.ncell, p.N
    ld = lambda n, shape: np.fromfile(tmp_path / n, np.float32).reshape(shape)
    g = host.Organoid(p)
    g.set_field(host.F_PHI, ld("ic_phi.f32", (nc, N, N)))
    g.set_field(host.F_RHO, ld("ic_rho.f32", (nc, N, N)))
    g.set_field(host.F_LUM, ld("ic_lum.f32", (N, N)))
    g.set_field(host.F_ECM, ld("ic_ecm.f32", (N, N)))
    g.set_field(host.F_HFUNC, np.zeros((nc, N, N), np.float32))
    g.prepare()
    g.relax_initial_state(0.001, 0.01)
    for name, f, shape in (("phi", host.F_PHI, (nc, N, N)), ("rho", host.F_RHO, (nc, N, N)),
                           ("lum", host.F_LUM, (N, N)), ("ecm", host.F_ECM, (N, N))):
        assert np.array_equal(ld(f"out_{name}.f32", shape), g.get_field(f)), name


REF_DRIVER = "/root/reference/main_code/run_with_active_gels.cpp"
REF_FUNCTIONS = "/root/reference/functions"


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="the reference sources are not on this machine")
def test_unmodified_reference_driver_links(tmp_path):
    """The reference's own driver (main_code/run_with_active_gels.cpp),
    unmodified and compiled where it lies with the reference's own headers,
    links against liborganoid_dropin.so / liborganoid_hip.so with the two
    one-line headers INTEGRATION.md tells a maintainer to add (cufft.h
    including organoid_dropin.h, an empty cublas.h) and a rendered
    constants.cpp: every function and table it references resolves in the
    drop-in, none in a stand-in."""
    (tmp_path / "cufft.h").write_text('/* signature types of the organoid drop-in library */\n'
                                      '#include "organoid_dropin.h"\n')
    (tmp_path / "cublas.h").write_text("/* unused by the reference's timestep */\n")
    obj = tmp_path / "driver.o"
    inc = ["-I", str(tmp_path), "-I", REF_FUNCTIONS, "-I", os.path.join(ROOT, "include")]
    subprocess.run(["g++", "-O1", "-std=c++14", "-Drestrict=__restrict", "-w"] + inc + ["-c", REF_DRIVER, "-o", str(obj)],
                   check=True)
    # every function and global the driver leaves undefined, except libc /
    # libstdc++ / libm and the constants.h globals (defined by constants.cpp)
    undef = set(re.findall(r"^\s+U (\S+)$", subprocess.run(["nm", "-C", str(obj)], check=True, capture_output=True,
                                                            text=True).stdout, re.M))
    consts = set(re.findall(r"^(?:const )?(?:int|float|org_real_t) (\w+)", open(os.path.join(
        ROOT, "tests", "dropin", "constants_two_cells.cpp")).read(), re.M))
    lib_syms = _exports(DROPIN, demangle=True)
    # the reference's tables (extern declarations of declare_tables.h and
    # declare_fft_tables.h, read as text) and its C++ functions
    tables = set()
    for h in ("declare_tables.h", "declare_fft_tables.h"):
        for decl in re.findall(r"^\s*extern\s+([^;]+);", open(os.path.join(REF_FUNCTIONS, h)).read(), re.M):
            tables.update(re.findall(r"(\w+)\s*(?:,|$)", decl))
    system = ("std::", "vtable for", "VTT for", "typeinfo", "_")
    ours = sorted(u for u in undef if not u.startswith(system) and u not in consts and ("(" in u or u in tables))
    assert len(ours) >= 30, ours           # the timestep's functions and the tables
    for u in ours:
        assert re.search(rf" [TBD] {re.escape(u)}$", lib_syms, re.M), u
    exe = tmp_path / "run_with_active_gels"
    libdir = os.path.join(PKG, "csrc")
    subprocess.run(["g++", "-o", str(exe), str(obj), os.path.join(ROOT, "tests", "dropin", "constants_two_cells.cpp"),
                    "-I", os.path.join(ROOT, "include"), "-L", os.path.join(libdir, "dropin"), "-lorganoid_dropin",
                    "-L", libdir, "-lorganoid_hip", "-Wl,--no-undefined"], check=True)
    assert exe.exists()


REF_HEADERS = ("constants.h", "declare_tables.h", "declare_fft_tables.h", "initialisation_functions.h",
               "useful_functions.h", "rho_functions.h", "cell_phase_functions.h")


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="the reference sources are not on this machine")
def test_retyped_reference_driver_links_f64(tmp_path):
    """The FP64 form of the reference: its driver and headers with every
    `float` read as `double` (a scratch copy under the test's temporary
    folder, nothing kept), compiled against the drop-in headers with
    ORGANOID_DROPIN_F64 set by the one-line cufft.h, links against
    liborganoid_dropin_f64.so with every symbol resolved."""
    src = tmp_path / "src"
    src.mkdir()
    retype = lambda text: re.sub(r"\bfloat\b", "double", text)  # noqa: E731
    for h in REF_HEADERS:
        (src / h).write_text(retype(open(os.path.join(REF_FUNCTIONS, h)).read()))
    (src / "run_with_active_gels.cpp").write_text(retype(open(REF_DRIVER).read()))
    (tmp_path / "cufft.h").write_text("/* signature types of the organoid drop-in library, FP64 form */\n"
                                      "#define ORGANOID_DROPIN_F64\n#include \"organoid_dropin.h\"\n")
    (tmp_path / "cublas.h").write_text("/* unused by the reference's timestep */\n")
    obj = tmp_path / "driver.o"
    inc = ["-I", str(tmp_path), "-I", str(src), "-I", os.path.join(ROOT, "include")]
    subprocess.run(["g++", "-O1", "-std=c++14", "-Drestrict=__restrict", "-w"] + inc +
                   ["-c", str(src / "run_with_active_gels.cpp"), "-o", str(obj)], check=True)
    undef = subprocess.run(["nm", "-C", str(obj)], check=True, capture_output=True, text=True).stdout
    assert "update_all_cortices(double*" in undef and "org_complex64*" in undef
    exe = tmp_path / "run_with_active_gels_f64"
    libdir = os.path.join(PKG, "csrc")
    subprocess.run(["g++", "-DORGANOID_DROPIN_F64", "-o", str(exe), str(obj),
                    os.path.join(ROOT, "tests", "dropin", "constants_two_cells.cpp"),
                    "-I", os.path.join(ROOT, "include"), "-L", os.path.join(libdir, "dropin"), "-lorganoid_dropin_f64",
                    "-L", libdir, "-lorganoid_hip", "-Wl,--no-undefined"], check=True)
    assert exe.exists()

"""Benchmark: cell-timesteps/s of the organoid timestep on one MI355X.

Workload: ``organoid64_1024`` -- 64 cells (radii U(16,24)) on a jittered hex
packing, 1024 x 1024 grid, zone_size 65 (FFT box 128), FP32 fields with FP64
reductions, inputs resident in HBM before the timed region.

A step is one full timestep (main_code/run_with_active_gels.cpp:118-184):
centroids, turnover labels, h, psi, areas, the spectral cortex update of
every cell, lumen/ECM update + reshaping, phase-field update + reshaping of
every cell, copy-back.  The timed form is the batched one (for all cells:
k_pre, then k_step + k_mid per step, k_step + k_post for the last step of a
replay; a replay of n steps is n / L hipGraphs of L steps plus one graph of
the n % L remaining steps, L autotuned).  The plain per-function drop-in
form (the reference's call sequence with its host round trips) is timed
beside it on a shorter run.

The timed window is steps [2 W, 2 W + K) of the trajectory (W warm-up steps,
the autotuner, which restores the state, W more warm-up steps).  In the
first few dozen steps the lumen/ECM reshaping loop runs many more sweeps
than later (its stop test is an absolute threshold on the whole grid's
change), so short windows cost more per step; the line reports the window
and the sweeps executed in it.

Timed beside it (``config``): the plain form, the uneven 8-organoid sweep
batch, the 256-cell 2048^2 workload, the FNO surrogate, the FP64 form on
the same workload (``fp64_form``) and the reference template's own
configuration (``template``: 6 cells, 1200^2, FFT box 398, on the
full-plane path in the float build's arithmetic) against the reference's
published steps/s for it.

Printed: one JSON line (see the repository's benchmark contract) with a
``roofline`` object for the dominant kernel (algorithmic bytes per launch /
HIP-event launch time on the library's stream, both over a replay of the
timed window) and a ``cpu_baseline`` object (the FP64 CPU oracle on a
bounded sample of the same workload, OpenMP threads).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-timesteps/s, 64 cells @1024^2: FP32+batch+FNO vs plain, phi L2err<=1e-3"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes of the dominant launch, k_step (FP32 = 4 B; M = FFT
# box side, W = window side, N = grid side; R_c, R_lum, R_ecm = reshaping
# sweeps executed):
#   spectral role, per cell:    read phi_n, phi_n-1, phi_n+1, rho_n, lum, ECM
#                               on the box, write rho_n on the box  -> 28 B x M^2
#   phase-field role, per cell: read phi_n, rho_n, hfunc_n, psi, h(lum), h(ECM)
#                               on the window, write phi_n and phin, read phi_n
#                               for the next centroid (support box >= window)
#                                                                   -> 36 B x W^2
#                               each sweep after the first: read phi, phi_old,
#                               write phi                      -> 12 B x W^2 (R_c - 1)
#   lumen/ECM role, per grid:   read lum, ECM, psi, write lumn, ECMn -> 20 B x N^2
#                               each sweep: read, read old, write  -> 12 B x N^2 (R_lum + R_ecm)
# These are the per-cell-step bytes B_alg of the reference's data flow
# restricted to windows and boxes, minus the psi pass (k_mid / k_pre).
def alg_bytes_k_step(nc: int, M: int, W: int, N: int, rc: float, rl: float, re: float) -> float:
    return nc * (28 * M * M + 36 * W * W + 12 * W * W * max(rc - 1.0, 0.0)) + N * N * (20 + 12 * (rl + re))


PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_step_organoid64_1024_pmc.md")


def surrogate_rates(host, org, nc: int, steps: int) -> dict:
    """The same organoid stepped with the FNO velocity surrogate (exact solve
    every 10th step, judge threshold off so every other step uses the
    network), and the network's forward pass alone.  Untrained weights: the
    time does not depend on them."""
    import torch

    from multicellular_phase_field_active_gel_organoids_amd import autotune, fno
    cfg = fno.FnoConfig(batch_max=nc)
    net = fno.FnoNet(cfg)
    net.set_params(fno.init_params(cfg, 0))
    tiles = autotune.tuned_fno(net, nc, reps=3, passes=1, train=False)
    org.step_surrogate(10, net=net, k_exact=10, threshold=1e30, it0=1)
    org.sync()
    t0 = time.perf_counter()
    exact, _ = org.step_surrogate(steps, net=net, k_exact=10, threshold=1e30, it0=1)
    org.sync()
    dt = time.perf_counter() - t0
    x = torch.rand((nc, 2, cfg.M, cfg.M), device="cuda")
    out = torch.empty_like(x)
    net.forward(x, out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        net.forward(x, out)
    e1.record()
    torch.cuda.synchronize()
    fwd = e0.elapsed_time(e1) / 10
    net.close()
    return {"cell_steps_per_s": round(nc * steps / dt, 2), "ms_per_step": round(1e3 * dt / steps, 4),
            "fno_forward_ms": round(fwd, 4), "fno_tiles": tiles["tiles"], "fno_tiles_source": tiles["source"],
            "exact_cell_solves": int(exact.sum()), "cell_steps": nc * steps,
            "form": "eager steps; FNO (4 layers, width 32, 12 modes, BF16 MFMA) for the velocity, exact solve "
                    "every 10th step; untrained weights (timing only)"}


def surrogate_spec(samples: int, train_steps: int, deploy_steps: int) -> int:
    """The surrogate at the specification's scale (a separate run, not the
    benchmark line): about `samples` per-cell training pairs (phi_n, rho_n) ->
    (vx, vy) on 128^2 boxes from the exact solve of the uneven 8-organoid
    sweep (steps 50.. of each member), training at batch 256 with Adam, the
    network's error on held-out states of organoid64_1024 (steps 100..105),
    then organoid64_1024 deployed for `deploy_steps` steps with the exact
    solve every 10th step and the spectral judge at 1e-3, against the exact
    steps of the same trajectory (and, at 10^4 steps, the FP64 oracle's
    fields from tests/golden/drift_organoid64_1024.npz when present).
    Prints one JSON line."""
    import numpy as np
    import torch

    from multicellular_phase_field_active_gel_organoids_amd import fno, host
    from multicellular_phase_field_active_gel_organoids_amd.configs import make_sweep, make_workload
    t_all = time.perf_counter()
    members = make_sweep()
    ncells = sum(w.params.ncell for w in members)
    per = max(1, -(-samples // ncells))
    xs, vs = [], []
    t0 = time.perf_counter()
    for w in members:
        g = host.Organoid(w.params, w.s0_cell)
        g.load_workload(w)
        g.step(50)
        x, v = g.surrogate_samples(per, it0=50)
        xs.append(x)
        vs.append(v)
        g.close()
    x = torch.cat(xs)
    v = torch.cat(vs)
    del xs, vs
    gen_s = time.perf_counter() - t0
    print(f"bench: {x.shape[0]} training pairs in {gen_s:.1f} s", file=sys.stderr, flush=True)
    cfg = fno.FnoConfig(batch_max=256)
    net = fno.FnoNet(cfg)
    net.set_params(fno.init_params(cfg, 7))
    sub = slice(None, None, max(1, x.shape[0] // 4096))
    net.set_scales(*fno.fit_scales(x[sub].cpu().numpy(), v[sub].cpu().numpy()))
    t0 = time.perf_counter()
    losses = fno.train(net, x, v, steps=train_steps, batch=256, lr=2e-3, lr_final=1e-4,
                       log=lambda m: print("bench: train " + m, file=sys.stderr, flush=True))
    train_s = time.perf_counter() - t0
    del x, v
    wl = make_workload("organoid64_1024")
    nc = wl.params.ncell
    g = host.Organoid(wl.params, wl.s0_cell)
    g.load_workload(wl)
    g.step(100)
    xv, vv = g.surrogate_samples(6, it0=100)
    g.close()
    pred = torch.cat([net.forward(xv[b:b + nc].contiguous()) for b in range(0, xv.shape[0], nc)])
    err = ((pred - vv).double().flatten(1).norm(dim=1) / vv.double().flatten(1).norm(dim=1)).cpu().numpy()
    del xv, vv, pred
    print(f"bench: held-out error median {np.median(err):.4f}", file=sys.stderr, flush=True)

    # deployment on organoid64_1024 from its initial state: exact every 10th
    # step, the spectral judge at 1e-3, against the exact trajectory
    ex_org = host.Organoid(wl.params, wl.s0_cell)
    ex_org.load_workload(wl)
    ex_org.set_option(host.OPT_WRITE_VELOCITY, 0)
    t0 = time.perf_counter()
    ex_org.step(deploy_steps, graph=True)
    ex_org.sync()
    exact_s = time.perf_counter() - t0
    su_org = host.Organoid(wl.params, wl.s0_cell)
    su_org.load_workload(wl)
    chunk = 1000
    exact_cells = 0
    rejected = 0
    judged = 0
    t0 = time.perf_counter()
    for it0 in range(0, deploy_steps, chunk):
        k = min(chunk, deploy_steps - it0)
        ex, res = su_org.step_surrogate(k, net=net, k_exact=10, threshold=1e-3, it0=it0)
        exact_cells += int(ex.sum())
        judged += int((res >= 0).sum())
        rejected += int((res > 1e-3).sum())
        print(f"bench: deployment step {it0 + k}: exact cell solves {exact_cells}", file=sys.stderr, flush=True)
    su_org.sync()
    sur_s = time.perf_counter() - t0
    net.close()
    fields = {}
    for name, f in (("phi", host.F_PHI), ("rho", host.F_RHO)):
        a = su_org.get_field(f).astype(np.float64)
        b_ = ex_org.get_field(f).astype(np.float64)
        fields[name] = {"vs_exact_gpu": float(np.linalg.norm(a - b_) / np.linalg.norm(b_))}
        fx_path = os.path.join(ROOT, "tests", "golden", "drift_organoid64_1024.npz")
        key = f"s{deploy_steps}_{name}"
        if os.path.exists(fx_path):
            fx = np.load(fx_path)
            if key + "_box" in fx:
                want = np.array(getattr(wl, name), dtype=np.float64, copy=True)
                for n, (i0, i1, j0, j1) in enumerate(fx[key + "_box"]):
                    want[n, i0:i1, j0:j1] = fx[f"{key}_{n}"]
                fields[name]["vs_fp64_oracle"] = float(np.linalg.norm(a - want) / np.linalg.norm(want))
                fields[name]["exact_gpu_vs_fp64_oracle"] = float(np.linalg.norm(b_ - want) / np.linalg.norm(want))
    ex_org.close()
    su_org.close()
    out = {"samples": int(per * ncells), "generation_s": round(gen_s, 1), "train_steps": train_steps, "batch": 256,
           "train_s": round(train_s, 1), "loss_first": round(float(losses[0]), 5),
           "loss_last_50": round(float(np.mean(losses[-50:])), 6),
           "heldout": "organoid64_1024 steps 100..105 (384 cells)",
           "heldout_rel_l2_median": round(float(np.median(err)), 5),
           "heldout_rel_l2_p90": round(float(np.percentile(err, 90)), 5),
           "deployment": {"workload": wl.name, "steps": deploy_steps, "k_exact": 10, "judge": "spectral",
                          "threshold": 1e-3, "cell_steps": nc * deploy_steps, "exact_cell_solves": exact_cells,
                          "judged": judged, "rejected_by_judge": rejected,
                          "ms_per_step": round(1e3 * sur_s / deploy_steps, 4),
                          "exact_ms_per_step": round(1e3 * exact_s / deploy_steps, 4), "fields": fields}}
    print(json.dumps({"surrogate_spec": out}), flush=True)
    print(f"bench: surrogate at spec {time.perf_counter() - t_all:.1f} s", file=sys.stderr, flush=True)
    return 0


def large_rates(host, make_workload, steps: int) -> dict:
    """organoid256_2048 (256 cells, 2048^2) in the batched form (8-step
    graphs, default launch shape) and in the plain per-function form."""
    wl = make_workload("organoid256_2048")
    p = wl.params
    nc = p.ncell
    org = host.Organoid(p, wl.s0_cell)
    org.load_workload(wl)
    org.set_option(host.OPT_WRITE_VELOCITY, 0)
    org.step(24, graph=True)
    org.sync()
    t0 = time.perf_counter()
    org.step(steps, graph=True)
    org.sync()
    dt = time.perf_counter() - t0
    org.close()
    plain = host.Organoid(p, wl.s0_cell)
    plain.load_workload(wl)
    plain.plain_step()
    plain.sync()
    t0 = time.perf_counter()
    for _ in range(3):
        plain.plain_step()
    plain.sync()
    pdt = time.perf_counter() - t0
    plain.close()
    rate, prate = nc * steps / dt, nc * 3 / pdt
    print(f"bench: organoid256_2048 {1e3 * dt / steps:.4f} ms/step", file=sys.stderr, flush=True)
    return {"workload": wl.name, "cells": nc, "grid": p.N, "steps": steps, "ms_per_step": round(1e3 * dt / steps, 4),
            "cell_steps_per_s": round(rate, 2), "plain_dropin_cell_steps_per_s": round(prate, 2),
            "speedup_vs_plain": round(rate / prate, 3)}


# The reference's own throughput figure: 100001 steps of its template
# configuration (6 cells, 1200^2, zone_size 200: FFT box 398) "runned about
# 25 times in 24 hours" on an unstated NVIDIA GPU, saving every 10000 steps
# (README.md:61): about 28.9 steps/s.
REFERENCE_TEMPLATE_STEPS_PER_S = 100001 * 25 / 86400.0


def template_rates(host, make_workload, steps: int) -> dict:
    """template6_1200 -- the reference template's own configuration -- on
    the full-plane path in FP32 (the reference's float arithmetic, bit for
    bit): the IC program's relaxation (org_relax_initial_state, untimed),
    20 warm-up steps, then `steps` timed steps."""
    wl = make_workload("template6_1200")
    p = wl.params
    org = host.Organoid(p, wl.s0_cell)
    org.load_workload(wl)
    t0 = time.perf_counter()
    org.relax_initial_state(0.001, 0.01)
    org.sync()
    relax_s = time.perf_counter() - t0
    org.prepare()
    org.step(20)
    org.sync()
    t0 = time.perf_counter()
    org.step(steps)
    org.sync()
    dt = time.perf_counter() - t0
    cells, lum, ecm = org.iterations()
    org.close()
    sps = steps / dt
    print(f"bench: template6_1200 {1e3 * dt / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return {"workload": wl.name, "cells": p.ncell, "grid": p.N, "zone_size": p.zone_size, "fft_box": p.M,
            "form": "full-plane path, FP32 (the float build's arithmetic), host-driven steps",
            "ic_relaxation_s": round(relax_s, 3), "timed_steps": [20, 20 + steps - 1],
            "ms_per_step": round(1e3 * dt / steps, 4), "steps_per_s": round(sps, 2),
            "cell_steps_per_s": round(p.ncell * sps, 2),
            "last_step_reshape_sweeps": {"cells": int(cells.max()), "lum": lum, "ecm": ecm},
            "reference_published_steps_per_s": round(REFERENCE_TEMPLATE_STEPS_PER_S, 2),
            "vs_reference_published": round(sps / REFERENCE_TEMPLATE_STEPS_PER_S, 3)}


def fp64_rates(host, wl, steps: int) -> dict:
    """The FP64 form (org_create_f64: the reference retyped to double, the
    full-plane path) on the bench workload, steps 5..5+steps-1."""
    p = wl.params
    org = host.Organoid(p, wl.s0_cell, precision="f64")
    org.load_workload(wl)
    org.step(5)
    org.sync()
    t0 = time.perf_counter()
    org.step(steps)
    org.sync()
    dt = time.perf_counter() - t0
    org.close()
    print(f"bench: FP64 form {1e3 * dt / steps:.3f} ms/step", file=sys.stderr, flush=True)
    return {"timed_steps": [5, 5 + steps - 1], "ms_per_step": round(1e3 * dt / steps, 4),
            "cell_steps_per_s": round(p.ncell * steps / dt, 2)}


def fno_forward_only(n: int) -> int:
    """n forward passes of the FNO surrogate on 64 random 128^2 crops (the
    organoid64_1024 batch), for kernel traces; prints the mean time."""
    import torch

    from multicellular_phase_field_active_gel_organoids_amd import fno
    cfg = fno.FnoConfig(batch_max=64)
    net = fno.FnoNet(cfg)
    net.set_params(fno.init_params(cfg, 0))
    x = torch.rand((64, 2, cfg.M, cfg.M), device="cuda")
    out = torch.empty_like(x)
    net.forward(x, out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(n):
        net.forward(x, out)
    e1.record()
    torch.cuda.synchronize()
    print(json.dumps({"fno_forward_ms": round(e0.elapsed_time(e1) / n, 4), "cells": 64, "passes": n}))
    net.close()
    return 0


def measured_traffic(path: str = PMC_SUMMARY):
    """HBM bytes per k_step launch read from the committed rocprofv3 --pmc
    summary (line 'traffic_bytes_per_launch: <int>'), or None.  It is the
    profile's figure for its own window (named in that file), not a count
    taken in this run."""
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        return None
    return None


def calib_copy(mib: int) -> int:
    """A plain copy kernel of known size (reads and writes `mib` MiB) for
    scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 2.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4}), flush=True)
    return 0


def sweep_rates(host, steps: int) -> dict:
    """The uneven 8-organoid sweep (configs.make_sweep): cell-timesteps/s of
    one batch (one launch per stage for all organoids) and of the same
    organoids stepped one after another, each with its own graphs."""
    from multicellular_phase_field_active_gel_organoids_amd.configs import make_sweep
    wls = make_sweep()

    def members():
        out = []
        for w in wls:
            m = host.Organoid(w.params, w.s0_cell)
            m.load_workload(w)
            m.set_option(host.OPT_WRITE_VELOCITY, 0)
            out.append(m)
        return out

    cells = sum(w.params.ncell for w in wls)
    ms = members()
    with host.Batch(ms) as b:
        b.step(16)
        b.sync()
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        tb = time.perf_counter() - t0
    for m in ms:
        m.close()
    ms = members()
    for m in ms:
        m.step(16)
        m.sync()
    t0 = time.perf_counter()
    for m in ms:
        m.step(steps)
    for m in ms:
        m.sync()
    ts = time.perf_counter() - t0
    for m in ms:
        m.close()
    print(f"bench: sweep8 batch {1e3 * tb / steps:.4f} ms/step, one by one {1e3 * ts / steps:.4f} ms/step",
          file=sys.stderr, flush=True)
    return {"organoids": len(wls), "cells": cells, "grids": [w.params.N for w in wls],
            "steps": steps, "batch_cell_steps_per_s": round(cells * steps / tb, 2),
            "batch_ms_per_step": round(1e3 * tb / steps, 5),
            "one_by_one_cell_steps_per_s": round(cells * steps / ts, 2)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="organoid64_1024")
    ap.add_argument("--plain-steps", type=int, default=40)
    ap.add_argument("--cpu-steps", type=int, default=8,
                    help="timesteps of the FP64 CPU oracle timed as cpu_baseline (~0.3-2.6 s each on 8 threads)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweep-steps", type=int, default=160,
                    help="steps of the uneven 8-organoid sweep batch (0: skip)")
    ap.add_argument("--large-steps", type=int, default=100,
                    help="steps of the 256-cell 2048^2 workload timed beside the main one (0: skip)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape instead of autotuning it first")
    ap.add_argument("--surrogate-steps", type=int, default=40,
                    help="steps timed with the FNO velocity surrogate (0: skip)")
    ap.add_argument("--template-steps", type=int, default=300,
                    help="steps of the reference template's configuration (6 cells, 1200^2, FFT box 398) on the "
                         "full-plane path (0: skip)")
    ap.add_argument("--fp64-steps", type=int, default=20,
                    help="steps of the bench workload in the FP64 form (0: skip)")
    ap.add_argument("--surrogate-spec", type=int, default=0, metavar="SAMPLES",
                    help="only run the surrogate at the specification's scale: SAMPLES training pairs (e.g. 200000), "
                         "batch 256, deployment on organoid64_1024 (see --spec-train-steps, --spec-deploy-steps)")
    ap.add_argument("--spec-train-steps", type=int, default=800)
    ap.add_argument("--spec-deploy-steps", type=int, default=10000)
    ap.add_argument("--fno-forward", type=int, default=0, metavar="N",
                    help="only run N forward passes of the FNO surrogate on 64 random crops (profiling)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        return calib_copy(args.calib_copy)
    if args.fno_forward:
        return fno_forward_only(args.fno_forward)
    if args.surrogate_spec:
        return surrogate_spec(args.surrogate_spec, args.spec_train_steps, args.spec_deploy_steps)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(see DESIGN.md), --gpus must be 1"}))
        return 2

    import numpy as np

    from multicellular_phase_field_active_gel_organoids_amd import host
    from multicellular_phase_field_active_gel_organoids_amd.configs import make_workload

    wl = make_workload(args.workload)
    p = wl.params
    nc = p.ncell
    print(f"bench: workload {wl.name} ncell={nc} N={p.N}", file=sys.stderr, flush=True)

    org = host.Organoid(p, wl.s0_cell)
    org.load_workload(wl)
    org.set_option(host.OPT_WRITE_VELOCITY, 0)

    # warmup (also builds the graph), the launch-shape autotuner (results do
    # not depend on it), warmup of the tuned graph, then the timed region
    org.step(max(args.warmup, 1), graph=True)
    org.sync()
    tuning = {"tuning": org.get_tuning(), "source": "default"}
    if not args.no_tune:
        from multicellular_phase_field_active_gel_organoids_amd import autotune
        r = autotune.tuned(org, steps=args.steps, warmup=0)     # on the steps about to be timed
        tuning = {"tuning": r["tuning"], "source": r["source"],
                  "tuned_on": f"the {args.steps} steps after the first warm-up (steps {max(args.warmup, 1)}.."
                              f"{max(args.warmup, 1) + args.steps - 1}), from the same state for every candidate"}
        if r["source"] == "tuned":
            tuning.update(default=r["default"], default_ms_per_step=round(r["default_ms"], 5),
                          tuned_ms_per_step=round(r["best_ms"], 5), candidates_timed=len(r["measured"]))
        print(f"bench: tuning {tuning}", file=sys.stderr, flush=True)
        org.step(max(args.warmup, 1), graph=True)
        org.sync()
    gs = org.get_tuning()["graph_steps"]
    w0 = 2 * max(args.warmup, 1)
    org.prepare_graph(args.steps)     # capture + instantiate the replayed graphs (set-up, state unchanged)
    org.sync()
    t0 = time.perf_counter()
    org.step(args.steps, graph=True, it0=w0)
    org.sync()
    dt = time.perf_counter() - t0
    value = nc * args.steps / dt
    ms_per_step = 1e3 * dt / args.steps
    print(f"bench: batched {ms_per_step:.4f} ms/step", file=sys.stderr, flush=True)
    form = (f"batched: k_step + k_mid per step for all cells (k_pre / k_post at the ends of a replay); "
            f"{args.steps // gs} hipGraph(s) of {gs} steps + {1 if args.steps % gs else 0} of "
            f"{args.steps % gs} steps, instantiated before the timed region; exact velocity solve (FNO surrogate "
            f"not used in the timed region; "
            f"phi error not checked in it: see tests/test_gpu_drift.py)")

    # the same window replayed eagerly, one step at a time, on a second
    # organoid: per-kernel times with HIP events on the library stream and
    # the reshaping sweeps of every step
    nprof = min(200, args.steps)
    rep_org = host.Organoid(p, wl.s0_cell)
    rep_org.load_workload(wl)
    rep_org.set_option(host.OPT_WRITE_VELOCITY, 0)
    rep_org.set_tuning(**org.get_tuning())
    rep_org.step(w0, graph=True)
    rep_org.sync()
    rep_org.profile(True)
    sweeps = {"cells_mean": 0.0, "lum": 0, "ecm": 0}
    for k in range(nprof):
        rep_org.step(1, graph=False, it0=w0 + k)
        ci, li, ei = rep_org.iterations()
        sweeps["cells_mean"] += float(np.mean(ci)) / nprof
        sweeps["lum"] += li
        sweeps["ecm"] += ei
    kt = rep_org.kernel_times()
    rep_org.profile(False)
    rep_org.close()
    per_kernel = {k: (ms / max(n, 1), n) for k, (ms, n) in kt.items()}
    avg_ms = per_kernel["step"][0]              # profiling slot of k_step
    bytes_launch = alg_bytes_k_step(nc, p.M, p.W, p.N, sweeps["cells_mean"], sweeps["lum"] / nprof,
                                    sweeps["ecm"] / nprof)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic() if wl.name == "organoid64_1024" else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_source": "profiles/k_step_organoid64_1024_pmc.md (rocprofv3 --pmc, its own window)"
                                  if traffic else None,
                "kernel": "k_step", "avg_launch_ms": round(avg_ms, 5),
                "alg_bytes_per_launch": int(bytes_launch),
                "window": f"steps {w0}..{w0 + nprof - 1}, eager replay, HIP events on the library stream"}
    # profiling slots of the batched step's launches (the other slots are
    # used by the per-function entry points)
    names = {"centres": "k_pre", "prep": "k_prep", "psi": "k_psi", "lumecm": "k_lumecm_update",
             "lumreshape": "k_lum_normals", "spectral": "k_spectral", "cells": "k_cells",
             "commit": "k_post", "step": "k_step", "mid": "k_mid"}
    kernel_ms = {names[k]: round(v[0], 5) for k, v in per_kernel.items() if v[1]}

    # plain drop-in form: the reference's per-function call sequence
    plain = host.Organoid(p, wl.s0_cell)
    plain.load_workload(wl)
    plain.plain_step()
    plain.sync()
    t0 = time.perf_counter()
    for _ in range(args.plain_steps):
        plain.plain_step()
    plain.sync()
    plain_rate = nc * args.plain_steps / (time.perf_counter() - t0) if args.plain_steps > 0 else None
    plain.close()

    sweep = None
    if args.sweep_steps > 0:
        sweep = sweep_rates(host, args.sweep_steps)

    large = None
    if args.large_steps > 0:
        large = large_rates(host, make_workload, args.large_steps)

    surrogate = None
    if args.surrogate_steps > 0:
        surrogate = surrogate_rates(host, org, nc, args.surrogate_steps)

    template = None
    if args.template_steps > 0:
        template = template_rates(host, make_workload, args.template_steps)

    fp64 = None
    if args.fp64_steps > 0:
        fp64 = fp64_rates(host, wl, args.fp64_steps)

    cpu = None
    if not args.no_cpu:
        from oracle.oracle import Oracle, threads
        o = Oracle(p, wl.s0_cell)
        o.load_workload(wl)
        t0 = time.perf_counter()
        for it in range(args.cpu_steps):
            o.step(it)
        cdt = time.perf_counter() - t0
        nthr = threads()
        cpu = {"value": round(nc * args.cpu_steps / cdt, 3), "unit": "cell-timesteps/s", "cores": nthr,
               "kind": "port",
               "sample": f"steps 0..{args.cpu_steps - 1} of {wl.name} ({nc} cells, {p.N}^2) on the FP64 CPU "
                         f"oracle (full-plane restatement of the reference, OpenMP over cells / rows), "
                         f"{nthr} threads, {cdt:.1f} s; these first steps run more lumen/ECM sweeps "
                         f"(281 at step 0) than the GPU's window"}
        o.close()

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "cell-timesteps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "cells": nc, "grid": p.N, "zone_size": p.zone_size,
                   "form": form,
                   "timed_steps": [w0, w0 + args.steps - 1],
                   "plain_dropin_cell_steps_per_s": round(plain_rate, 2) if plain_rate else None,
                   "speedup_vs_plain": round(value / plain_rate, 3) if plain_rate else None,
                   "kernel_ms": kernel_ms,
                   "reshape_sweeps_in_window": {"steps": nprof, "cells_mean_per_step": round(sweeps["cells_mean"], 3),
                                                "lum_total": sweeps["lum"], "ecm_total": sweeps["ecm"]},
                   "launch_tuning": tuning, "sweep8": sweep, "large": large, "surrogate": surrogate,
                   "fp64_form": fp64, "template": template},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    org.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
